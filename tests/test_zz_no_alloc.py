"""oakg_local_analysis_csr_noalloc (the entry point that allocates no device
memory: its buffers come from a fixed scratch array per GPU) against
oakg_local_analysis_csr on the same device inputs: two calls in a row, the
second smaller, then a third after a call of the allocating entry point;
sizes beyond the scratch are rejected before any device work.  Needs an
MI355X."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _problem(seed, nz, zsize, m, N):
    g = np.random.default_rng(seed)
    sizes = g.integers(1, zsize + 1, nz)
    zone_ptr = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    n = int(zone_ptr[-1])
    Sf = np.asfortranarray(g.standard_normal((n, N)) / np.sqrt(N - 1))
    HSf = np.asfortranarray(g.standard_normal((m, N)) / np.sqrt(N - 1))
    xf = g.standard_normal(n); Hxf = g.standard_normal(m)
    yo = Hxf + 0.3 * g.standard_normal(m)
    r2 = g.uniform(0.05, 0.5, m)
    sp, si, sw = [0], [], []
    for z in range(nz):
        cnt = int(g.integers(0, min(m, 3 * N) + 1)) if z % 7 else 0
        idx = np.sort(g.choice(m, cnt, replace=False))
        si.extend(idx.tolist()); sw.extend(g.uniform(0.05, 1.0, cnt).tolist())
        sp.append(sp[-1] + cnt)
    return dict(zone_ptr=zone_ptr, sel_ptr=np.array(sp, np.int64), sel_idx=np.array(si, np.int32),
                sel_w=np.array(sw), xf=xf, Hxf=Hxf, yo=yo, r2=r2, Sf=Sf, HSf=HSf, n=n, N=N, m=m,
                nz=nz)


def _call(fn, ctx, d):
    import torch
    from oak_amd import _lib
    dev = torch.device("cuda:0")
    t = lambda a, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(a)).to(dev, dt)
    n, N, m, nz = d["n"], d["N"], d["m"], d["nz"]
    keep = [t(d["zone_ptr"], torch.int64), t(d["sel_ptr"], torch.int64),
            t(d["sel_idx"], torch.int32), t(d["sel_w"]), t(d["xf"]), t(d["Hxf"]), t(d["yo"]),
            t(d["r2"]), t(d["Sf"].T).t(), t(d["HSf"].T).t()]
    zp, sp, si, sw, xf, Hxf, yo, r2, Sf, HSf = keep
    xa = torch.full((n,), np.nan, dtype=torch.float64, device=dev)
    Sa = torch.full((N, n), np.nan, dtype=torch.float64, device=dev).t()
    o = _lib.Obs(m, None, None, None, None, yo.data_ptr(), r2.data_ptr(), None)
    rc = fn(ctx.handle, nz, zp.data_ptr(), sp.data_ptr(), si.data_ptr(), sw.data_ptr(),
            C.byref(o), N, xf.data_ptr(), Hxf.data_ptr(), Sf.data_ptr(), n, HSf.data_ptr(), m,
            xa.data_ptr(), Sa.data_ptr(), n, None, 1, _lib.OAKG_DEVICE_PTRS)
    torch.cuda.synchronize()
    return rc, xa.cpu().numpy(), Sa.cpu().numpy()


@pytest.mark.parametrize("N", [24, 128])
def test_noalloc_matches_allocating_entry_point(N):
    from oak_amd import _lib
    L = _lib.load()
    ctx = _lib.Context(0)
    try:
        big = _problem(1 + N, 700, 40, 3000, N)
        small = _problem(2 + N, 90, 9, 400, N)
        for d in (big, small):
            rc0, xa0, Sa0 = _call(L.oakg_local_analysis_csr, ctx, d)
            assert rc0 == 0, ctx.last_error()
            rc1, xa1, Sa1 = _call(L.oakg_local_analysis_csr_noalloc, ctx, d)
            assert rc1 == 0, ctx.last_error()
            np.testing.assert_array_equal(xa1, xa0)
            np.testing.assert_array_equal(Sa1, Sa0)
            assert not np.array_equal(xa1, d["xf"])
        # the allocating entry point in between, then the big case again
        rc2, xa2, Sa2 = _call(L.oakg_local_analysis_csr_noalloc, ctx, big)
        rc3, xa3, Sa3 = _call(L.oakg_local_analysis_csr, ctx, big)
        assert rc2 == 0 and rc3 == 0
        np.testing.assert_array_equal(xa2, xa3)
        np.testing.assert_array_equal(Sa2, Sa3)
    finally:
        ctx.close()


def test_noalloc_rejects_sizes_beyond_the_scratch():
    """m so large that HS alone exceeds the scratch: OAKG_EARG with a
    message, before any device work (the arrays are never read)."""
    from oak_amd import _lib
    L = _lib.load()
    ctx = _lib.Context(0)
    try:
        m = (_lib.OAKG_NOALLOC_SCRATCH_BYTES // (8 * 130)) + 1000
        # non-NULL addresses that are never read: the size check comes first
        dummy = np.zeros(4)
        a = dummy.ctypes.data
        o = _lib.Obs(m, None, None, None, None, a, a, None)
        rc = L.oakg_local_analysis_csr_noalloc(ctx.handle, 1, a, a, a, a, C.byref(o), 128, a, a,
                                               a, 1, a, m, a, a, 1, None, 1,
                                               _lib.OAKG_DEVICE_PTRS)
        assert rc == _lib.OAKG_EARG
        assert "scratch" in ctx.last_error()
        rc = L.oakg_local_analysis_csr_noalloc(ctx.handle, 1, a, a, a, a, C.byref(o), 128, a, a,
                                               a, 1, a, m, a, a, 1, None, 1, 0)
        assert rc == _lib.OAKG_EARG and "device pointers" in ctx.last_error()
    finally:
        ctx.close()
