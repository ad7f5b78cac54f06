"""The zone loop captured in a HIP graph and replayed (OAKG_TUNE_GRAPH), and
the tuning knobs (chunks in flight, zones per QL wave): every replay and
every setting gives the bits of the eager, default run; a replay reads the
current contents of the inputs; eager work between the replays does not
disturb them.  Needs an MI355X."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def problem():
    from oak_amd import synth
    from oak_amd import assimilation as A
    P = synth.column_ocean(64, 64, 1, 40, 3000, 100, seed_cfg=700, hlen=4, vlen=1, L=3.0,
                           cutoff=9.0)
    assert P.nzones >= 4096  # several zone chunks on several streams
    dev = torch.device("cuda:0")
    t = lambda a, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(a)).to(dev, dt)
    g = A.Geometry(t(P.zone_ptr, torch.int64), t(P.zpos), t(P.corr_len), t(P.max_len), t(P.ox),
                   t(P.oy))
    rp, co, va = A.coo_to_csr(P.m, P.Hi, P.Hj, P.Hs)
    H = (t(rp, torch.int64), t(co, torch.int32), t(va))
    return P, g, H, t(P.Hshift), t(P.yo), t(P.rmse2), t(P.E.T).t()


def _run(ctx, pr, yo=None, out=None):
    from oak_amd import assimilation as A
    P, g, H, hs, yo0, r2, E = pr
    Ea = out if out is not None else torch.empty((E.shape[1], E.shape[0]), dtype=E.dtype,
                                                 device=E.device).t()
    A.ensemble_analysis(g, H, hs, yo0 if yo is None else yo, r2, E, ctx=ctx, out=Ea)
    torch.cuda.synchronize()
    return Ea.cpu().numpy()


def test_graph_replays_match_eager(problem):
    from oak_amd._lib import Context
    ref_ctx = Context(0)
    ref = _run(ref_ctx, problem)
    yo2 = problem[4] + 0.05
    ref2 = _run(ref_ctx, problem, yo=yo2)
    ref_ctx.close()
    ctx = Context(0)
    ctx.set_tuning(graph=1)
    E = problem[6]
    out = torch.empty((E.shape[1], E.shape[0]), dtype=torch.float64, device=E.device).t()
    junk = torch.randn(2048, 2048, device=E.device, dtype=torch.float64)
    for rep in range(4):
        got = _run(ctx, problem, out=out)
        np.testing.assert_array_equal(got, ref)
        junk = junk @ junk.T / 2048.0  # eager work between the replays
        torch.cuda.synchronize()
    tu = ctx.tuning()
    assert tu["graph_captures"] == 1 and tu["graph_replays"] == 3, tu
    # the same graph with new observation values (same buffers)
    yo_buf = problem[4].clone()
    yo_buf.copy_(yo2)
    got2 = _run(ctx, problem, yo=yo_buf, out=out)
    np.testing.assert_array_equal(got2, ref2)
    ctx.close()


@pytest.mark.parametrize("slots,ql,order", [(1, 8, 0), (2, 32, 1), (4, 16, 0), (3, 16, 1),
                                           (1, 8, 1), (3, 4, 1)])
def test_tuning_knobs_same_bits(problem, slots, ql, order):
    from oak_amd._lib import Context
    ref_ctx = Context(0)
    ref = _run(ref_ctx, problem)
    ref_ctx.close()
    ctx = Context(0)
    ctx.set_tuning(slots=slots, ql_zones=ql, order=order)
    np.testing.assert_array_equal(_run(ctx, problem), ref)
    ctx.close()


def test_order_with_graph_and_timing(problem):
    """The ordered loop under graph replay and with phase timing (one chunk in
    order on the caller's stream) gives the same bits."""
    from oak_amd._lib import Context
    ref_ctx = Context(0)
    ref = _run(ref_ctx, problem)
    ref_ctx.close()
    ctx = Context(0)
    ctx.set_tuning(order=1, graph=1)
    for _ in range(3):
        np.testing.assert_array_equal(_run(ctx, problem), ref)
    assert ctx.tuning()["order"] == 1
    ctx.set_tuning(graph=0)
    ctx.set_timing(True)
    np.testing.assert_array_equal(_run(ctx, problem), ref)
    ctx.close()


@pytest.fixture(scope="module")
def problem128():
    from oak_amd import synth
    from oak_amd import assimilation as A
    P = synth.column_ocean(24, 20, 2, 100, 1500, 60, seed_cfg=701, hlen=4, vlen=1, L=3.0,
                           cutoff=9.0)
    dev = torch.device("cuda:0")
    t = lambda a, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(a)).to(dev, dt)
    g = A.Geometry(t(P.zone_ptr, torch.int64), t(P.zpos), t(P.corr_len), t(P.max_len), t(P.ox),
                   t(P.oy))
    rp, co, va = A.coo_to_csr(P.m, P.Hi, P.Hj, P.Hs)
    H = (t(rp, torch.int64), t(co, torch.int32), t(va))
    return P, g, H, t(P.Hshift), t(P.yo), t(P.rmse2), t(P.E.T).t()


@pytest.mark.parametrize("tri_split,order,graph", [(1, 0, 0), (1, 1, 0), (1, 1, 1), (0, 1, 1)])
def test_two_launch_reduction_same_bits(problem128, tri_split, order, graph):
    """N = 100 (NP = 128): the Householder reduction as two kernels, with and
    without the zone order and graph replay, gives the bits of the default
    run (one kernel, zones in their order, eager launches)."""
    from oak_amd._lib import Context
    assert problem128[6].shape[1] == 100
    ref_ctx = Context(0)
    ref = _run(ref_ctx, problem128)
    ref_ctx.close()
    ctx = Context(0)
    ctx.set_tuning(tri_split=tri_split, order=order, graph=graph)
    for _ in range(2):
        np.testing.assert_array_equal(_run(ctx, problem128), ref)
    assert ctx.tuning()["tri_split"] == tri_split
    ctx.close()


@pytest.fixture(scope="module")
def problem128_big():
    from oak_amd import synth
    from oak_amd import assimilation as A
    P = synth.column_ocean(64, 64, 1, 100, 2500, 90, seed_cfg=702, hlen=4, vlen=1, L=3.0,
                           cutoff=8.0)
    assert P.nzones >= 4096  # several zone chunks on several streams
    dev = torch.device("cuda:0")
    t = lambda a, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(a)).to(dev, dt)
    g = A.Geometry(t(P.zone_ptr, torch.int64), t(P.zpos), t(P.corr_len), t(P.max_len), t(P.ox),
                   t(P.oy))
    rp, co, va = A.coo_to_csr(P.m, P.Hi, P.Hj, P.Hs)
    H = (t(rp, torch.int64), t(co, torch.int32), t(va))
    return P, g, H, t(P.Hshift), t(P.yo), t(P.rmse2), t(P.E.T).t()


@pytest.mark.parametrize("persist,order,graph,slots,dual",
                         [(1, 0, 0, 3, 0), (2, 1, 1, 3, 0), (1, 1, 1, 1, 0), (2, 0, 0, 2, 0),
                          (1, 1, 0, 3, 64), (2, 1, 1, 3, 64)])
def test_persistent_gram_same_bits(problem128_big, persist, order, graph, slots, dual):
    """N = 100: the Gram as a persistent kernel (a zone per workgroup, or each
    wave the next zone of its quarter of the tiles, from counters cleared by a
    fill kernel in every chunk, also under graph replay) gives the bits of the
    one-workgroup-per-zone kernel with the same other settings, in FP64 and
    with FP32 anomalies."""
    from oak_amd._lib import Context
    pr = problem128_big
    E = pr[6]
    pr32 = pr[:6] + (E.float(),)
    for prob in (pr, pr32):
        ref_ctx = Context(0)
        ref_ctx.set_tuning(dual=dual, order=order, slots=slots)
        ref = _run(ref_ctx, prob)
        ref_ctx.close()
        ctx = Context(0)
        ctx.set_tuning(persist=persist, order=order, graph=graph, slots=slots, dual=dual)
        for _ in range(3 if graph else 1):
            np.testing.assert_array_equal(_run(ctx, prob), ref)
        tu = ctx.tuning()
        assert tu["persist"] == persist
        if graph:
            assert tu["graph_replays"] == 2, tu
        ctx.close()


def test_tuning_rejects_bad_values():
    from oak_amd import OakError
    from oak_amd._lib import Context
    ctx = Context(0)
    for kw in (dict(slots=0), dict(slots=5), dict(ql_zones=12), dict(graph=2), dict(order=2),
               dict(tri_split=5), dict(dual=48), dict(persist=3), dict(persist=-1)):
        with pytest.raises(OakError):
            ctx.set_tuning(**kw)
    ctx.close()
