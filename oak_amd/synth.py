"""Synthetic inputs of the benchmark configurations (seeded, reproducible).

Ensembles are Gaussian random fields with covariance exp(-r^2/l^2): white
noise filtered by a separable Gaussian kernel of standard deviation l/2 per
axis (truncated at 4 sigma) and rescaled to unit variance.  Observations are
point values of a truth member through bilinear (2-D) or trilinear (3-D)
interpolation plus Gaussian noise of the stated error.  Random numbers come
from numpy's Philox generator seeded with 20260000 + configuration index.

Layouts follow OAK: state vectors column-major per variable (i fastest), then
permuted so that each zone is contiguous (assimilation.F90:869-880); the
observation operator is a COO list of (obs, permuted state index, weight).
"""
from dataclasses import dataclass, field

import numpy as np

SEED0 = 20260000


def rng(cfg):
    return np.random.Generator(np.random.Philox(SEED0 + cfg))


def gaussian_kernel(length):
    s = length / 2.0
    r = int(np.ceil(4 * s))
    x = np.arange(-r, r + 1, dtype=np.float64)
    k = np.exp(-0.5 * (x / s) ** 2)
    return k / np.sqrt(np.sum(k * k))


def _filter_axis(a, k, axis):
    """Periodic-free convolution along one axis (zero padded, 'same' size)."""
    a = np.moveaxis(a, axis, -1)
    r = (len(k) - 1) // 2
    pad = [(0, 0)] * (a.ndim - 1) + [(r, r)]
    ap = np.pad(a, pad)
    out = np.zeros_like(a)
    for i, kv in enumerate(k):
        out += kv * ap[..., i:i + a.shape[-1]]
    return np.moveaxis(out, -1, axis)


def grf(g, shape, lengths, nmembers):
    """nmembers fields of the given shape; lengths per axis (None = white)."""
    out = np.empty((nmembers,) + tuple(shape))
    for k in range(nmembers):
        pad = [int(np.ceil(2 * l)) if l else 0 for l in lengths]
        big = tuple(s + 2 * p for s, p in zip(shape, pad))
        w = g.standard_normal(big)
        for ax, l in enumerate(lengths):
            if l:
                w = _filter_axis(w, gaussian_kernel(l), ax)
        sl = tuple(slice(p, p + s) for p, s in zip(pad, shape))
        f = w[sl]
        out[k] = f / f.std()
    return out


@dataclass
class Problem:
    """One analysis problem in OAK's (zone-permuted) layout."""
    name: str
    n: int
    N: int
    m: int
    zone_ptr: np.ndarray
    zpos: np.ndarray          # nz x 4
    corr_len: np.ndarray
    max_len: np.ndarray
    E: np.ndarray             # n x N, Fortran order
    Hi: np.ndarray            # COO rows
    Hj: np.ndarray            # COO cols (permuted state index)
    Hs: np.ndarray
    Hshift: np.ndarray
    ox: np.ndarray
    oy: np.ndarray
    yo: np.ndarray
    rmse2: np.ndarray
    metric: int = 0
    loctype: int = 1
    truth: np.ndarray = None
    info: dict = field(default_factory=dict)

    @property
    def nzones(self):
        return len(self.zone_ptr) - 1


def bilinear(x, y, nx, ny):
    """(i0, j0, weights[4]) of the 1-based cell coordinates x, y."""
    i0 = np.clip(np.floor(x).astype(np.int64), 1, nx - 1)
    j0 = np.clip(np.floor(y).astype(np.int64), 1, ny - 1)
    a = x - i0
    b = y - j0
    w = np.stack([(1 - a) * (1 - b), a * (1 - b), (1 - a) * b, a * b], 1)
    return i0, j0, w


def config0(N=16, seed_cfg=0):
    """64x64 grid, one variable (n = 4096), N = 16 GRF members (l = 8 cells),
    400 observations uniform in [1,64)^2, bilinear H, sigma = 0.1, taper L = 5,
    cutoff 15, one zone per grid point."""
    g = rng(seed_cfg)
    nx = ny = 64
    n = nx * ny
    fields = grf(g, (nx, ny), (8, 8), N + 1)
    E = np.asfortranarray(np.stack([f.ravel(order="F") for f in fields[:N]], 1))
    truth = fields[N].ravel(order="F")
    m = 400
    x = g.uniform(1, 64, m)
    y = g.uniform(1, 64, m)
    i0, j0, w = bilinear(x, y, nx, ny)
    Hi = np.repeat(np.arange(m), 4)
    cols = []
    for di, dj in ((0, 0), (1, 0), (0, 1), (1, 1)):
        cols.append((i0 + di - 1) + nx * (j0 + dj - 1))
    Hj = np.stack(cols, 1).ravel()
    Hs = w.ravel()
    sigma = 0.1
    Hd_truth = np.zeros(m)
    np.add.at(Hd_truth, Hi, Hs * truth[Hj])
    yo = Hd_truth + sigma * g.standard_normal(m)
    gi, gj = np.meshgrid(np.arange(1, nx + 1, dtype=float), np.arange(1, ny + 1, dtype=float),
                         indexing="ij")
    zpos = np.zeros((n, 4))
    zpos[:, 0] = gi.ravel(order="F")
    zpos[:, 1] = gj.ravel(order="F")
    return Problem("cfg0", n, N, m, np.arange(n + 1, dtype=np.int64), zpos, np.full(n, 5.0),
                   np.full(n, 15.0), E, Hi.astype(np.int32), Hj.astype(np.int32), Hs,
                   np.zeros(m), x, y, yo, np.full(m, sigma ** 2), truth=truth)


def column_ocean(nx, ny, nz_lev, N, n_ssh, n_prof_cols, seed_cfg, hlen=10.0, vlen=5.0,
                 L=6.0, cutoff=18.0, sig_ssh=0.03, sig_t=0.5, ssh_density=None):
    """Ocean-like layout of the real-shape configurations, at any size:
    variables T, S, u, v on nx x ny x nz_lev plus 2-D SSH; one zone per
    water column (4*nz_lev + 1 elements, first element T at level 1).
    SSH observations along straight tracks (or drawn from ssh_density),
    T profiles at random columns, one value per level."""
    g = rng(seed_cfg)
    ncol = nx * ny
    s = 4 * nz_lev + 1
    n = ncol * s
    # state: zone z = column (i, j) with i fastest; inside a zone T(1..K), S, u, v, SSH
    E = np.empty((n, N), order="F")
    truth = None
    ssh_base = 4 * nz_lev
    for k in range(N + 1):
        f3 = grf(g, (nx, ny, nz_lev), (hlen, hlen, vlen), 4)
        f2 = grf(g, (nx, ny), (hlen, hlen), 1)[0]
        col = np.empty((ncol, s))
        for v in range(4):
            col[:, v * nz_lev:(v + 1) * nz_lev] = f3[v].reshape(ncol, nz_lev, order="F")
        col[:, ssh_base] = f2.ravel(order="F")
        if k < N:
            E[:, k] = col.ravel()
        else:
            truth = col.ravel()
    P = column_ocean_obs(g, nx, ny, nz_lev, truth, n_ssh, n_prof_cols, L, cutoff, sig_ssh,
                         sig_t, ssh_density)
    P.E = E
    P.N = N
    return P


def lognormal_density(median_count, sigma_log, cutoff, corr_cells=60.0):
    """Observation positions from a smooth log-normal density field chosen so
    that the count within the cutoff disc has roughly the given median."""
    def draw(g, nx, ny):
        f = grf(g, (nx, ny), (corr_cells, corr_cells), 1)[0]
        area = np.pi * cutoff ** 2
        dens = (median_count / area) * np.exp(sigma_log * f)
        cnt = g.poisson(dens)
        ii, jj = np.nonzero(cnt)
        reps = cnt[ii, jj]
        x = np.repeat(ii + 1.0, reps) + g.uniform(0, 1, reps.sum())
        y = np.repeat(jj + 1.0, reps) + g.uniform(0, 1, reps.sum())
        ok = (x < nx) & (y < ny)
        return x[ok], y[ok]
    return draw


def calibrated_swaths(median=150.0, sigma_log=0.97, cutoff=18.0, n_prof=60, nz_lev=30,
                      field_len=50.0, gap_len=80.0, gap_frac=0.12, bins=1000, iters=80):
    """Observation positions whose local counts (observations within the
    cutoff of each column) follow a log-normal distribution of the given
    median and log standard deviation, with gaps of no data.

    SSH observations are drawn from a cell density exp(T(u)) on the columns,
    u the rank (0..1) of a smooth random field of correlation length
    field_len: dense swaths where the field is high, sparse data where it is
    low.  Cells where a second smooth field (gap_len) is in its
    lowest gap_frac hold no data (neither SSH nor profiles), so the columns
    deep inside these gaps have no local observation.  n_prof profile
    columns of nz_lev values are drawn first, outside the gaps.  T is
    calibrated by quantile matching: the expected local count of each column
    (the density summed over the cutoff disc, by FFT convolution, plus the
    profile values within the cutoff) is compared with the log-normal
    quantile of its rank among the columns with data, and T is corrected per
    rank bin of u, until the expected counts follow the log-normal; a final
    scale puts their median on target.  Returns a callable for
    column_ocean_obs(ssh_density=...) giving (sx, sy, px, py)."""
    def draw(g, nx, ny):
        from scipy.signal import fftconvolve
        from scipy.stats import norm
        f = grf(g, (nx, ny), (field_len, field_len), 1)[0]
        h = grf(g, (nx, ny), (gap_len, gap_len), 1)[0]
        mask = (h > np.quantile(h, gap_frac)).astype(np.float64)
        # profiles: columns outside the gaps
        px, py = [], []
        while len(px) < n_prof:
            x, y = g.uniform(1, nx), g.uniform(1, ny)
            if mask[int(x) - 1, int(y) - 1] > 0:
                px.append(x); py.append(y)
        px = np.array(px); py = np.array(py)
        ii, jj = np.meshgrid(np.arange(1, nx + 1, dtype=float), np.arange(1, ny + 1, dtype=float),
                             indexing="ij")
        prof = np.zeros((nx, ny))
        for x, y in zip(px, py):
            prof += nz_lev * ((ii - x) ** 2 + (jj - y) ** 2 <= cutoff ** 2)
        R = int(np.ceil(cutoff)) + 1
        yy, xx = np.mgrid[-R:R + 1, -R:R + 1]
        disc = (xx ** 2 + yy ** 2 <= cutoff ** 2).astype(np.float64)
        u = (np.argsort(np.argsort(f.ravel())).reshape(f.shape) + 0.5) / f.size
        b = np.minimum((u * bins).astype(np.int64), bins - 1)
        logT = np.log(median / disc.sum()) + sigma_log * norm.ppf((np.arange(bins) + 0.5) / bins)

        def expected(logT):
            return fftconvolve(np.exp(logT[b]) * mask, disc, mode="same") + prof
        for _ in range(iters):
            ec = expected(logT)
            ok = ec > 0.5
            rk = np.argsort(np.argsort(ec[ok]))
            tgt = np.log(median) + sigma_log * norm.ppf((rk + 0.5) / ok.sum())
            lr = np.zeros_like(ec)
            lr[ok] = tgt - np.log(ec[ok])
            s = np.bincount(b[ok], weights=lr[ok], minlength=bins)
            c = np.bincount(b[ok], minlength=bins)
            logT += 0.7 * np.where(c > 0, s / np.maximum(c, 1), 0.0)
        for _ in range(20):  # the median of the columns with data on target
            ec = expected(logT)
            logT += np.log(median / np.median(ec[ec > 0.5]))
        cnt = g.poisson(np.exp(logT[b]) * mask)
        ci, cj = np.nonzero(cnt)
        reps = cnt[ci, cj]
        x = np.repeat(ci + 1.0, reps) + g.uniform(0, 1, reps.sum())
        y = np.repeat(cj + 1.0, reps) + g.uniform(0, 1, reps.sum())
        ok = (x < nx) & (y < ny)
        return x[ok], y[ok], px, py
    return draw


def dense_H(P):
    """Dense m x n observation operator (small problems only)."""
    H = np.zeros((P.m, P.n))
    np.add.at(H, (P.Hi, P.Hj), P.Hs)
    return H


# ---------------------------------------------------------------------------
# Large configurations: ensembles generated on the GPU with torch (seeded),
# observations and H on the host.
# ---------------------------------------------------------------------------
def _grf_torch(gen, shape, lengths, device, dtype):
    import torch
    pads = [int(np.ceil(2 * l)) if l else 0 for l in lengths]
    big = [s + 2 * p for s, p in zip(shape, pads)]
    w = torch.randn(big, generator=gen, device=device, dtype=torch.float64)
    for ax, l in enumerate(lengths):
        if not l:
            continue
        k = gaussian_kernel(l)
        r = (len(k) - 1) // 2
        x = w.movedim(ax, -1)
        xp = torch.nn.functional.pad(x, (r, r))
        out = torch.zeros_like(x)
        for i, kv in enumerate(k):
            out.add_(xp[..., i:i + x.shape[-1]], alpha=float(kv))
        w = out.movedim(-1, ax)
    sl = tuple(slice(p, p + s) for p, s in zip(pads, shape))
    f = w[sl]
    return (f / f.std()).to(dtype)


@dataclass
class DeviceProblem:
    """Problem with the ensemble resident on the GPU (column-major n x N)."""
    host: Problem            # geometry, observations, H (host arrays); host.E is None
    E: object                # torch tensor (n, N), stride (1, n)


def column_ocean_gpu(nx, ny, nz_lev, N, cfg, n_ssh=150000, n_prof_cols=1700, hlen=10.0,
                     vlen=5.0, L=6.0, cutoff=18.0, sig_ssh=0.03, sig_t=0.5, ssh_density=None,
                     dtype=None, device="cuda:0"):
    """The ocean-column layout of column_ocean() with the members drawn on the GPU."""
    import torch
    dtype = dtype or torch.float64
    g = rng(cfg)
    gen = torch.Generator(device=device)
    gen.manual_seed(SEED0 + cfg)
    ncol = nx * ny
    s = 4 * nz_lev + 1
    n = ncol * s
    E = torch.empty((N, n), dtype=dtype, device=device)
    truth = None
    import sys
    import time
    t0 = time.time()
    for k in range(N + 1):
        if k % 16 == 0:
            print("# synth: member %d/%d (%.1f s)" % (k, N + 1, time.time() - t0), file=sys.stderr,
                  flush=True)
        col = torch.empty((ncol, s), dtype=dtype, device=device)
        for v in range(4):
            f = _grf_torch(gen, (nx, ny, nz_lev), (hlen, hlen, vlen), device, dtype)
            col[:, v * nz_lev:(v + 1) * nz_lev] = f.permute(1, 0, 2).reshape(ncol, nz_lev)
        f2 = _grf_torch(gen, (nx, ny), (hlen, hlen), device, dtype)
        col[:, 4 * nz_lev] = f2.t().reshape(ncol)
        if k < N:
            E[k] = col.reshape(n)
        else:
            truth = col.reshape(n).double().cpu().numpy()
    P = column_ocean_obs(g, nx, ny, nz_lev, truth, n_ssh, n_prof_cols, L, cutoff, sig_ssh, sig_t,
                         ssh_density)
    P.N = N
    return DeviceProblem(P, E.t())


def column_ocean_obs(g, nx, ny, nz_lev, truth, n_ssh, n_prof_cols, L, cutoff, sig_ssh, sig_t,
                     ssh_density=None):
    """Observations, H and geometry of the ocean-column layout (host)."""
    ncol = nx * ny
    s = 4 * nz_lev + 1
    n = ncol * s
    ssh_base = 4 * nz_lev
    if ssh_density is None:
        ntr = max(1, n_ssh // 1000)
        per = int(np.ceil(n_ssh / ntr))
        xs, ys = [], []
        while sum(len(x) for x in xs) < n_ssh:
            x0, y0 = g.uniform(1, nx), g.uniform(1, ny)
            ang = g.uniform(0, np.pi)
            t = np.linspace(-1.5, 1.5, per) * max(nx, ny)
            xx = x0 + t * np.cos(ang)
            yy = y0 + t * np.sin(ang)
            ok = (xx >= 1) & (xx < nx) & (yy >= 1) & (yy < ny)
            xs.append(xx[ok]); ys.append(yy[ok])
        sx = np.concatenate(xs)[:n_ssh]
        sy = np.concatenate(ys)[:n_ssh]
    else:
        sxy = ssh_density(g, nx, ny)
        sx, sy = sxy[0], sxy[1]
    nss = len(sx)
    i0, j0, w = bilinear(sx, sy, nx, ny)
    Hi = [np.repeat(np.arange(nss), 4)]
    cols = []
    for di, dj in ((0, 0), (1, 0), (0, 1), (1, 1)):
        zc = (i0 + di - 1) + nx * (j0 + dj - 1)
        cols.append(zc * s + ssh_base)
    Hj = [np.stack(cols, 1).ravel()]
    Hs = [w.ravel()]
    if ssh_density is not None and len(sxy) == 4:
        # the density drew the profile columns as well
        px, py = sxy[2], sxy[3]
        n_prof_cols = len(px)
    else:
        px = g.uniform(1, nx, n_prof_cols)
        py = g.uniform(1, ny, n_prof_cols)
    pi0, pj0, pw = bilinear(px, py, nx, ny)
    npro = n_prof_cols * nz_lev
    Hi.append(np.repeat(nss + np.arange(npro), 4))
    pc = np.stack([(pi0 + di - 1) + nx * (pj0 + dj - 1)
                   for di, dj in ((0, 0), (1, 0), (0, 1), (1, 1))], 1)
    lev = np.arange(nz_lev)
    Hj.append((pc[:, None, :] * s + lev[None, :, None]).reshape(-1))
    Hs.append(np.repeat(pw[:, None, :], nz_lev, axis=1).reshape(-1))
    Hi = np.concatenate(Hi).astype(np.int32)
    Hj = np.concatenate(Hj).astype(np.int32)
    Hs = np.concatenate(Hs)
    m = nss + npro
    ox = np.concatenate([sx, np.repeat(px, nz_lev)])
    oy = np.concatenate([sy, np.repeat(py, nz_lev)])
    sig = np.concatenate([np.full(nss, sig_ssh), np.full(npro, sig_t)])
    Htruth = np.zeros(m)
    np.add.at(Htruth, Hi, Hs * truth[Hj])
    yo = Htruth + sig * g.standard_normal(m)
    gi, gj = np.meshgrid(np.arange(1, nx + 1, dtype=float), np.arange(1, ny + 1, dtype=float),
                         indexing="ij")
    zpos = np.zeros((ncol, 4))
    zpos[:, 0] = gi.ravel(order="F")
    zpos[:, 1] = gj.ravel(order="F")
    zone_ptr = np.arange(ncol + 1, dtype=np.int64) * s
    return Problem("ocean%dx%dx%d" % (nx, ny, nz_lev), n, 0, m, zone_ptr, zpos,
                   np.full(ncol, L), np.full(ncol, cutoff), None, Hi, Hj, Hs, np.zeros(m), ox, oy,
                   yo, sig ** 2, truth=truth)


def config1_gpu(device="cuda:0"):
    """360x180x30 x {T,S,u,v} + SSH (n = 7,840,800), N = 64, 150k SSH track
    observations + 1,700 T profiles x 30 levels, L = 6, cutoff 18, FP64."""
    return column_ocean_gpu(360, 180, 30, 64, 1, device=device)


def config2_gpu(N=128, device="cuda:0", dtype=None):
    """config 1's grid with N = 128; SSH observations in dense swaths and
    sparse regions, 60 T profiles of 30 levels, and gaps without data, drawn
    so that the local counts within the 18-cell cutoff are log-normal with
    median 150 and log standard deviation 0.97 (calibrated_swaths): about 5 %
    of the columns have no observation and the largest counts reach ~3000."""
    return column_ocean_gpu(360, 180, 30, N, 2, n_prof_cols=60,
                            ssh_density=calibrated_swaths(150.0, 0.97, 18.0, n_prof=60,
                                                          nz_lev=30),
                            dtype=dtype, device=device)


def config4_gpu(scale=1, device="cuda:0", dtype=None):
    """720x360x40 x {T,S,u,v} + SSH (n = 41,731,200 at scale 1), N = 256 FP32
    members; SSH observations in swaths and 10,000 T profiles of 40 levels
    (about 1.4M + 0.4M values), localisation L = 3 cells with cutoff 9 (3 L),
    the swaths drawn so that the local counts are log-normal with median
    1900 and log standard deviation 0.45, with gaps of no data.  scale > 1
    divides both horizontal sizes and the number of profiles (scale^2) and
    keeps the local counts."""
    import torch
    nx, ny = 720 // scale, 360 // scale
    n_prof = max(1, 10000 // (scale * scale))
    return column_ocean_gpu(nx, ny, 40, 256, 4, n_prof_cols=n_prof, L=3.0, cutoff=9.0,
                            ssh_density=calibrated_swaths(1900.0, 0.45, 9.0, n_prof=n_prof,
                                                          nz_lev=40),
                            dtype=dtype or torch.float32, device=device)


def config3(nx=256, N=32, m=4000, seed_cfg=3):
    """Differentiable-analysis configuration: the config 0 layout on an
    nx x nx grid (n = nx^2, one zone per point), N GRF members (l = 8 cells),
    m observations uniform in [1, nx)^2 with bilinear H, taper L = 5, cutoff 15.
    The observation errors are heterogeneous (sigma = 0.05 for even, 0.2 for
    odd observation indices): a per-observation log sigma is worth tuning."""
    g = rng(seed_cfg)
    ny = nx
    n = nx * ny
    fields = grf(g, (nx, ny), (8, 8), N + 1)
    E = np.asfortranarray(np.stack([f.ravel(order="F") for f in fields[:N]], 1))
    truth = fields[N].ravel(order="F")
    x = g.uniform(1, nx, m)
    y = g.uniform(1, ny, m)
    i0, j0, w = bilinear(x, y, nx, ny)
    Hi = np.repeat(np.arange(m), 4)
    cols = []
    for di, dj in ((0, 0), (1, 0), (0, 1), (1, 1)):
        cols.append((i0 + di - 1) + nx * (j0 + dj - 1))
    Hj = np.stack(cols, 1).ravel()
    Hs = w.ravel()
    sigma = np.where(np.arange(m) % 2 == 0, 0.05, 0.2)
    Hd_truth = np.zeros(m)
    np.add.at(Hd_truth, Hi, Hs * truth[Hj])
    yo = Hd_truth + sigma * g.standard_normal(m)
    gi, gj = np.meshgrid(np.arange(1, nx + 1, dtype=float), np.arange(1, ny + 1, dtype=float),
                         indexing="ij")
    zpos = np.zeros((n, 4))
    zpos[:, 0] = gi.ravel(order="F")
    zpos[:, 1] = gj.ravel(order="F")
    return Problem("cfg3", n, N, m, np.arange(n + 1, dtype=np.int64), zpos, np.full(n, 5.0),
                   np.full(n, 15.0), E, Hi.astype(np.int32), Hj.astype(np.int32), Hs,
                   np.zeros(m), x, y, yo, sigma ** 2, truth=truth)
