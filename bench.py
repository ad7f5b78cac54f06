"""Benchmark of the local ensemble square-root analysis on one MI355X.

One step = one call of the analysis (oak_assim's ensemble path: H·E, anomalies,
selection, per-zone S^T R^-1 S, eigensolve, transform, Ea) over the whole
synthetic ensemble, with every input already resident in HBM.

    python bench.py [--steps K] [--warmup W] [--config 2] [--precision both|fp64|mixed]

The default workload is config 2 (360x180x30x4+SSH grid, N=128, clustered
observations): the mixed-precision analysis (FP32 state, FP32 MFMA
contractions, FP64 eigensolve) is the headline ``value``; the FP64 analysis of
the same inputs is timed beside it (``fp64``, ``mixed_speedup_vs_fp64``) and
the mixed result is compared with it (``mixed_err_vs_fp64``).

Prints one JSON line (rank 0).  The metric is analysed state elements per
second: n / t_step, an element being one entry of xa with its N-member row of
Ea.  ``roofline`` is computed for the dominant kernel from device-event times
recorded inside the timed steps; ``cpu_baseline`` times the oracle (OAK's
algorithm restated in C, LAPACK dsyev + BLAS dgemm, OpenMP over zones) on a
contiguous sample of zones of the same workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64 = 78.6e12   # MI355X FP64 (vector and matrix), FLOP/s
PEAK_HBM = 8.0e12     # MI355X HBM3E, B/s


def build_problem(cfg, device, scale=1):
    from oak_amd import synth
    if cfg == 0:
        import torch
        P = synth.config0()
        E = torch.as_tensor(np.ascontiguousarray(P.E.T)).to(device).t()
        return synth.DeviceProblem(P, E), "64x64 grid, N=16, 400 obs"
    if cfg == 1:
        return synth.config1_gpu(device), "360x180x30x4+SSH (n=7,840,800), N=64, 201k obs"
    if cfg == 2:
        return synth.config2_gpu(device=device), "360x180x30x4+SSH (n=7,840,800), N=128, clustered obs"
    if cfg == 4:
        nx, ny = 720 // scale, 360 // scale
        return (synth.config4_gpu(scale, device=device),
                "%dx%dx40x4+SSH (n=%d), N=256 FP32 members, swath + profile obs, L=3, cutoff 9"
                % (nx, ny, nx * ny * 161))
    raise SystemExit("unknown config %d" % cfg)


def device_inputs(DP, device):
    import torch
    from oak_amd.assimilation import Geometry, coo_to_csr
    P = DP.host
    t = lambda a, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(a)).to(device, dt)
    g = Geometry(t(P.zone_ptr, torch.int64), t(P.zpos), t(P.corr_len), t(P.max_len), t(P.ox),
                 t(P.oy), metric=P.metric, loctype=P.loctype)
    rp, co, va = coo_to_csr(P.m, P.Hi, P.Hj, P.Hs)
    H = (t(rp, torch.int64), t(co, torch.int32), t(va))
    return g, H, t(P.Hshift), t(P.yo), t(P.rmse2)


def flops_model(N, counts, zone_sizes, dual_lt=0):
    """Algorithmic FLOP per analysis, per phase, from the per-zone local
    observation counts p_z (zones with p_z > 0), counted as the kernels
    compute them; with the dual bucket on (N > 64, 1 <= p_z < dual_lt) those
    zones count the p x p path instead of the Gram and the eigensolve:
      gram   PD (PD + 1) N            (G = S~ S~^T, PD = 64 padded rows)
      eig    9 PD^3 + 2 N PD p + 4 N p (the 64 x 64 eigensolve and the lift
                                       W = S~^T U c, a, M^1/2 1)
    and otherwise:
      gram   N (N + 1) p + 2 N p      (S^T R^-1 S, a symmetric rank-p update
                                       counted as BLAS dsyrk is, and
                                       S^T R^-1 d)
      eig    9 N^3                    (symmetric eigensolve with vectors,
                                       Golub & Van Loan's count for
                                       tridiagonalisation + QL)
      trans  N^2 (N + 1) + 10 N^2     (T = Y Y^T Omega: the symmetric
                                       product Y Y^T and Omega applied as its
                                       rank-structured correction; OAK's dense
                                       form, RotateVector + three N x N GEMMs,
                                       would be 6 N^3)
      update 2 s N^2 + 2 s N          (Sf_z [T | a])"""
    act = counts > 0
    dual = act & (counts < (dual_lt if N > 64 else 0))
    prim = act & ~dual
    p = counts[prim].astype(np.float64)
    pd = counts[dual].astype(np.float64)
    s = zone_sizes[act].astype(np.float64)
    nact = float(act.sum())
    PD = 64.0
    return {
        "gram": float(np.sum(N * (N + 1.0) * p + 2.0 * N * p)) + PD * (PD + 1.0) * N * len(pd),
        "eig": 9.0 * N ** 3 * float(prim.sum())
               + float(np.sum(9.0 * PD ** 3 + 2.0 * N * PD * pd + 4.0 * N * pd)),
        "trans": (N * N * (N + 1.0) + 10.0 * N * N) * nact,
        "update": float(np.sum(2.0 * s * N * N + 2.0 * s * N)),
    }


def cpu_baseline(DP, target_s=15.0, gpu_results=None):
    """Oracle (LAPACK backend, OpenMP dynamic over zones, OAK's brute-force
    selection) on a strided sample of zones spread over the whole domain (every
    k-th zone, so that dense and sparse observation regions are represented in
    proportion); the sampled zones' state rows are gathered into one slice.

    gpu_results: {label: Ea (device tensor, n x N)} of the GPU analyses of the
    same inputs (inflation 1, no clamp): their rows of the sampled zones are
    compared with the oracle's analysis, Ea = xa + sqrt(N-1) Sa
    (assimilation.F90:3318-3320), by oak_amd.errors.ensemble_errors.
    Returns (baseline dict, {label: errors})."""
    import torch
    from oak_amd.errors import ensemble_errors
    from oracle import oracle as O
    P = DP.host
    E = DP.E
    N = E.shape[1]
    lapack = O.load_lapack()
    # the host cores of one GPU's share of the box (16 per GPU); OMP_NUM_THREADS
    # overrides
    cores = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(os.cpu_count() or 1, 16)
    # obs-space anomalies for every observation (computed on the GPU: this
    # leg times the zone loop of locAnalysis, not H·E)
    rows = torch.as_tensor(P.Hi.astype(np.int64), device=E.device)
    cols = torch.as_tensor(P.Hj.astype(np.int64), device=E.device)
    vals = torch.as_tensor(P.Hs, device=E.device)
    HE = torch.zeros((P.m, N), dtype=torch.float64, device=E.device)
    HE.index_add_(0, rows, vals[:, None] * E[cols, :].double())
    Hxf = HE.mean(dim=1)
    HSf = ((HE - Hxf[:, None]) / np.sqrt(N - 1)).cpu().numpy()
    Hxf = Hxf.cpu().numpy()
    nz = P.nzones
    zs = np.diff(P.zone_ptr)

    def run(stride, sel=None):
        if sel is None:
            sel = np.arange(stride // 2, nz, stride)
        starts = P.zone_ptr[sel]
        rows_ = np.concatenate([np.arange(s, s + c) for s, c in zip(starts, zs[sel])])
        Ez = E[torch.as_tensor(rows_, device=E.device), :].double().cpu().numpy()
        xf = Ez.mean(axis=1)
        Sf = np.asfortranarray((Ez - xf[:, None]) / np.sqrt(N - 1))
        zp = np.concatenate([[0], np.cumsum(zs[sel])]).astype(np.int64)
        geo = dict(zone_ptr=zp, zpos=np.ascontiguousarray(P.zpos[sel]),
                   L=np.ascontiguousarray(P.corr_len[sel]),
                   cutoff=np.ascontiguousarray(P.max_len[sel]), metric=P.metric,
                   loctype=P.loctype, ox=P.ox, oy=P.oy)
        t0 = time.perf_counter()
        xa, Sa, _ = O.loc_analysis_geo(geo, xf, Hxf, P.yo, P.rmse2, Sf, HSf, nthreads=cores)
        return time.perf_counter() - t0, len(rows_), len(sel), (rows_, Ez, xa, Sa)

    probe_stride = max(1, nz // max(4 * cores, 64))
    t, ne, nzs, _ = run(probe_stride)
    per_zone = t / nzs
    print("# cpu baseline probe: %d zones in %.2f s" % (nzs, t), file=sys.stderr, flush=True)
    want = max(4 * cores, int(target_s / max(per_zone, 1e-6)))
    stride = max(1, nz // want)
    t, ne, nzs, (rows_, Ez, xa, Sa) = run(stride)
    base = {"value": ne / t, "unit": "analysed state elems/s", "cores": cores, "kind": "port",
            "sample": "%d zones (every %d-th of %d; %d state elements) of the same workload, "
                      "zone loop of locAnalysis with OAK's brute-force selection; %s "
                      "eigensolver, %.1f s" % (nzs, stride, nz, ne,
                                               "LAPACK dsyev (OpenBLAS)" if lapack else "Jacobi", t)}
    errs = {}
    if gpu_results:
        Eo = xa[:, None] + np.sqrt(N - 1.0) * Sa
        del Sa
        ri = torch.as_tensor(rows_, device=E.device)
        sel_all = np.arange(stride // 2, nz, stride)
        zone_of_row = np.repeat(np.arange(len(sel_all)), zs[sel_all])
        for label, Ea in gpu_results.items():
            Eg = Ea[ri, :].double().cpu().numpy()
            errs[label] = dict(ensemble_errors(Eg, Eo, Ez), zones=int(nzs),
                               elements=int(ne))
            if label == "fp64" and lapack:
                errs[label]["floor"] = solver_floor(run, Eg, Eo, Ez, zone_of_row, sel_all, O)
            del Eg
            print("# %s vs oracle on the sample: %s" % (label, errs[label]), file=sys.stderr,
                  flush=True)
    return base, errs


def solver_floor(run, Eg, Eo, Ez, zone_of_row, sel_all, O, nworst=64):
    """How far two FP64 CPU analyses of the same zones (the oracle with LAPACK
    dsyev, as OAK, and with its own cyclic Jacobi eigensolver) are apart, on
    the zones where the GPU's FP64 analysis is furthest from the LAPACK one;
    every measure is relative to the whole sample's increment and anomaly
    (as ensemble_errors), so it reads against the same bound."""
    xf = Ez.mean(axis=1)
    xo = Eo.mean(axis=1)
    inc = np.abs(xo - xf).max()
    ar = np.abs(Eo - xo[:, None]).max()
    dg = np.abs((Eg - Eg.mean(axis=1)[:, None]) - (Eo - xo[:, None])).max(axis=1)
    per_zone = np.zeros(len(sel_all))
    np.maximum.at(per_zone, zone_of_row, dg)
    worst = np.sort(np.argsort(per_zone)[-nworst:])
    O.set_backend(0)
    try:
        _, _, _, (rows_j, Ezj, xaj, Saj) = run(1, sel=sel_all[worst])
    finally:
        O.set_backend(1)
    Ej = xaj[:, None] + np.sqrt(Ezj.shape[1] - 1.0) * Saj
    pick = np.isin(zone_of_row, worst)
    Eo_w, Eg_w = Eo[pick], Eg[pick]

    def rel(A, B):
        a, b = A.mean(axis=1), B.mean(axis=1)
        return {"mean_err": float(np.abs(a - b).max() / inc),
                "spread_err": float(np.abs((A - a[:, None]) - (B - b[:, None])).max() / ar)}
    return {"zones": int(len(worst)), "oracle_lapack_vs_oracle_jacobi": rel(Ej, Eo_w),
            "gpu_vs_oracle_lapack": rel(Eg_w, Eo_w)}


PEAK_FP32_MFMA = 157.3e12  # MI355X dense FP32 matrix, FLOP/s


def run_analysis(args, ctx, g, H, Hshift, yo, rmse2, E, Ea, label):
    """Warmup + timed steps of one precision; returns (ms/step, phase ms, stats).

    The timed steps run as a caller would run them (no per-phase events, so the
    zone chunks overlap on two streams); the per-phase device times come from
    separate passes with phase timing on, where the chunks run in order."""
    import torch
    from oak_amd import autotune
    from oak_amd.assimilation import ensemble_analysis
    N = E.shape[1]
    if args.tune:
        # time every launch shape on this workload; the table entry is kept
        # in oak_amd/tuning.json
        run = lambda: ensemble_analysis(g, H, Hshift, yo, rmse2, E, ctx=ctx, out=Ea)
        best, times = autotune.tune(run, ctx, space=autotune.space_for(N),
                                    log=lambda s: print(s, file=sys.stderr, flush=True))
        autotune.save_entry(N, ctx.precision, best, times, "config%d" % args.config)
        print("# %s autotuned: %s" % (label, best), file=sys.stderr, flush=True)
    settings = autotune.apply(ctx, N, ctx.precision)
    if args.serial:
        ctx.set_tuning(graph=0)
    print("# %s launch settings: %s" % (label, settings), file=sys.stderr, flush=True)
    ctx.set_timing(args.serial)
    for i in range(args.warmup):
        ensemble_analysis(g, H, Hshift, yo, rmse2, E, ctx=ctx, out=Ea)
        torch.cuda.synchronize()
        print("# %s warmup %d done: %s" % (label, i, ctx.stats()), file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ensemble_analysis(g, H, Hshift, yo, rmse2, E, ctx=ctx, out=Ea)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    st = ctx.stats()
    print("# %s timed %d steps: %.3f ms/step %s %s" % (label, args.steps, ms, st, ctx.tuning()),
          file=sys.stderr, flush=True)
    st = dict(st, tuning=ctx.tuning())
    ctx.set_timing(True)
    phase = {"gram": 0.0, "eig": 0.0, "tmat": 0.0, "update": 0.0, "select": 0.0, "prep": 0.0}
    npass = max(1, min(args.steps, 2))
    t1 = time.perf_counter()
    for _ in range(npass):
        ensemble_analysis(g, H, Hshift, yo, rmse2, E, ctx=ctx, out=Ea)
        stp = ctx.stats()  # phase times of this call (device events)
        for k in phase:
            phase[k] += stp["ms_" + k]
    torch.cuda.synchronize()
    for k in phase:
        phase[k] /= npass
    print("# %s phase pass (chunks in order): %.3f ms/step %s" % (
        label, (time.perf_counter() - t1) * 1e3 / npass, {k: round(v, 2) for k, v in phase.items()}),
        file=sys.stderr, flush=True)
    ctx.set_timing(False)
    return ms, phase, st


def run_report(args, H, Hshift, yo, rmse2, Ef, Ea):
    """Assim's forecast and analysis reports (assimilation.F90:3389-3391) of
    the FP64 ensembles, by oakg_ensemble_report: H·E + Hshift, observation
    anomalies and the statistics, timed per call after one warm-up."""
    import torch
    from oak_amd import Context, OAKG_FP64
    from oak_amd.assimilation import REPORT_FIELDS, ensemble_report
    ctx = Context(0, OAKG_FP64)
    E = Ef.double() if Ef.dtype != torch.float64 else Ef
    ensemble_report(H, Hshift, yo, rmse2, E, ctx=ctx)
    torch.cuda.synchronize()
    k = max(1, args.steps)
    t0 = time.perf_counter()
    for _ in range(k):
        rf = ensemble_report(H, Hshift, yo, rmse2, E, ctx=ctx)
        ra = ensemble_report(H, Hshift, yo, rmse2, Ea, ctx=ctx)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / (2 * k)
    ctx.close()
    rf, ra = rf.cpu().numpy()[0], ra.cpu().numpy()[0]
    print("# report: %.3f ms per call; forecast %s; analysis %s" % (
        ms, {k: round(float(v), 6) for k, v in zip(REPORT_FIELDS, rf)},
        {k: round(float(v), 6) for k, v in zip(REPORT_FIELDS, ra)}),
        file=sys.stderr, flush=True)
    return {"ms_per_call": ms, "forecast": dict(zip(REPORT_FIELDS, rf.tolist())),
            "analysis": dict(zip(REPORT_FIELDS, ra.tolist()))}


def phase_rooflines(F, phase, mixed, N=128):
    """Achieved rate of every phase against the MI355X peak of the arithmetic
    it runs on: the Gram (FP64 MFMA accumulation in both paths), the
    eigensolve (FP64), the transform and update (FP32 MFMA on the mixed path,
    FP64 otherwise)."""
    work = {"gram": F["gram"], "eig": F["eig"], "tmat": F["trans"], "update": F["update"]}
    if phase["tmat"] == 0.0:  # N > 64: transform and update in one kernel
        work["update"] += work["tmat"]
        del work["tmat"]
    out = {}
    for k, w in work.items():
        if phase[k] <= 0:
            continue
        # N > 128: the transform and update run on FP64 MFMA in both paths
        peak = PEAK_FP32_MFMA if (mixed and N <= 128 and k in ("update", "tmat")) else PEAK_FP64
        ach = w / (phase[k] * 1e-3)
        out[k] = {"ms": round(phase[k], 3), "flop": w, "achieved": ach / 1e12,
                  "peak": peak / 1e12, "unit": "TFLOP/s", "frac": ach / peak}
    return out


def roofline(F, phase, mixed, eig_name, config, N):
    """Dominant phase of the step and its achieved rate against the MI355X peak
    of the arithmetic it runs on (FP32 MFMA for the mixed contractions, FP64
    otherwise)."""
    work = {"gram": F["gram"], "eig": F["eig"], "tmat": F["trans"], "update": F["update"]}
    # N > 64: transform and update run in one kernel, booked as the update
    fused = phase["tmat"] == 0.0
    if fused:
        work["update"] += work["tmat"]
    kern = max(work, key=lambda k: phase[k])
    gk = "k_gram256" if N > 128 else ("k_gram128" if N > 64 else "k_gram")
    names = {
        "gram": ("%s<float> (S^T R^-1 S from FP32 S, v_mfma_f64_16x16x4f64)" % gk if mixed else
                 "%s (S^T R^-1 S, v_mfma_f64_16x16x4f64)" % gk),
        "eig": eig_name,
        "tmat": "k_tmat (T = Y Y^T Omega, v_mfma_f64_16x16x4f64)",
        "update": ("k_update_f32 (Ea = xa + A' W, v_mfma_f32_16x16x4f32)" if mixed else
                   "k_update (Ea = xa + A' W, v_mfma_f64_16x16x4f64)"),
    }
    if N > 128:
        names["tmat"] = "k_tmat256 (T = Y Y^T Omega, v_mfma_f64_16x16x4f64)"
        names["update"] = ("k_update256 (Ea = xa + A' T from FP32 members on "
                           "v_mfma_f64_16x16x4f64)" if mixed else
                           "k_update256 (Ea = xa + A' T, v_mfma_f64_16x16x4f64)")
    if fused:
        names["update"] = ("k_xform128f (T = Y Y^T Omega in LDS on v_mfma_f32_16x16x4f32, then "
                           "Ea = xa + A' T on v_mfma_f32_16x16x4f32)" if mixed else
                           "k_xform128 (T = Y Y^T Omega in LDS on v_mfma_f64_16x16x4f64, then "
                           "Ea = xa + A' T on v_mfma_f64_16x16x4f64)")
    peak = PEAK_FP32_MFMA if (mixed and N <= 128 and kern == "update") else PEAK_FP64
    ach = work[kern] / (phase[kern] * 1e-3)
    return {"bound": "mfma", "kernel": names[kern], "achieved": ach / 1e12, "peak": peak / 1e12,
            "unit": "TFLOP/s", "frac": ach / peak, "traffic": phase_traffic(kern, mixed, config)}


# the kernels of each phase at N > 64 (the observation-space bucket's
# kernels and the Jacobi fallback passes included)
PHASE_KERNELS = {"eig": ["k_tridiag", "k_tridiag_lower", "k_formq128", "k_tql", "k_tql_apply128",
                         "k_eig128", "k_tql_apply", "k_tql_apply64", "k_formq64", "k_eigvec_epi",
                         "k_eig", "k_dual_lift"],
                 "gram": ["k_gram128", "k_gram128_persist", "k_dual_gram"],
                 "update": ["k_xform128", "k_xform128f"]}


def phase_traffic(kern, mixed, config):
    """HBM bytes moved by the phase's kernels in one analysis of the given
    precision, from the calibrated rocprofv3 FETCH_SIZE/WRITE_SIZE runs
    committed under profiles/ (per call, every dispatch of a kernel in the
    call summed; config 2 only, None otherwise)."""
    path = os.path.join(ROOT, "profiles", "traffic_config%d.json" % config)
    if not os.path.exists(path):
        return None
    tab = json.load(open(path)).get("bytes_per_call", {}).get("mixed" if mixed else "fp64")
    names = PHASE_KERNELS.get(kern)
    if not tab or not names:
        return None
    return float(sum(tab[k]["read"] + tab[k]["write"] for k in names if k in tab))


def run_config3(args, device):
    """Differentiable analysis (config 3: 256 x 256 grid, N = 32, 4000 obs):
    the innovation likelihood and its adjoint gradient (oakg_local_nll) timed
    per evaluation, checked against central differences (h = 1e-6), then 50
    Adam steps tuning per-observation log sigma, L and the inflation."""
    import torch
    from oak_amd import synth
    from oak_amd.nll import evaluate, problem_from_ensemble, tune
    P = synth.config3()
    prob = problem_from_ensemble(P, device=device)
    nz, m = P.nzones, P.m
    ls0 = torch.full((m,), float(np.log(0.3)), dtype=torch.float64, device=device)
    L0, a0 = 2.5, 1.0
    s2 = torch.exp(2 * ls0)
    corr = torch.full((nz,), L0, dtype=torch.float64, device=device)
    for _ in range(args.warmup):
        evaluate(prob, s2, corr, a0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        J, gth, gLz, ga = evaluate(prob, s2, corr, a0)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    print("# config3: %.3f ms per likelihood + gradient" % ms, file=sys.stderr, flush=True)
    # central differences of the GPU value
    h = 1e-6
    f = lambda ls_, L_, a_: float(evaluate(prob, torch.exp(2 * ls_),
                                           torch.full((nz,), L_, dtype=torch.float64,
                                                      device=device), a_)[0])
    fd_err = {}
    gmax = float(gth.abs().max())
    worst = 0.0
    for l in (0, 1, 1999, 3998):
        e = torch.zeros_like(ls0); e[l] = h
        fd = (f(ls0 + e, L0, a0) - f(ls0 - e, L0, a0)) / (2 * h)
        worst = max(worst, abs(fd - float(gth[l])) / max(1.0, gmax))
    fd_err["log_sigma"] = worst
    fdL = (f(ls0, L0 + h, a0) - f(ls0, L0 - h, a0)) / (2 * h)
    fd_err["L"] = abs(fdL - float(gLz.sum())) / max(1.0, abs(fdL))
    fda = (f(ls0, L0, a0 + h) - f(ls0, L0, a0 - h)) / (2 * h)
    fd_err["inflation"] = abs(fda - float(ga)) / max(1.0, abs(fda))
    print("# config3: central-difference errors %s" % fd_err, file=sys.stderr, flush=True)
    t1 = time.perf_counter()
    ls, L1, a1, hist = tune(prob, ls0, L0, a0, steps=50, lr=0.05)
    torch.cuda.synchronize()
    t_adam = time.perf_counter() - t1
    sig = torch.exp(ls).cpu().numpy()
    # the same with one sigma per class of observations (even / odd indices,
    # generated with 0.05 / 0.2): what one innovation vector can identify
    cls = np.arange(m) % 2
    lsc, L2, a2, hist2 = tune(prob, torch.full((2,), float(np.log(0.3)), dtype=torch.float64),
                              L0, a0, steps=200, lr=0.05, classes=cls)
    sigc = torch.exp(lsc).cpu().numpy()
    print("# config3: tied sigma per class %.4f / %.4f (generated 0.05 / 0.2), L %.2f, "
          "inflation %.3f" % (sigc[0], sigc[1], L2, a2), file=sys.stderr, flush=True)
    out = {
        "metric": "analysed state elems/s",
        "value": P.n / (ms * 1e-3),
        "unit": "analysed state elems/s (likelihood + adjoint gradient per evaluation)",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded GRF ensemble, truth member, heterogeneous obs errors)",
        "config": {"workload": "config3: 256x256 grid, N=32, 4000 obs, innovation NLL + adjoint",
                   "n": P.n, "N": P.N, "m": m, "zones": nz},
        "fd_check": {"h": h, "rel_err": fd_err},
        "adam": {"steps": 50, "lr": 0.05, "seconds": t_adam, "J_first": hist[0],
                 "J_last": hist[-1], "L": [L0, L1], "inflation": [a0, a1],
                 "sigma_start": 0.3,
                 "sigma_mean_true_0.05": float(sig[0::2].mean()),
                 "sigma_mean_true_0.2": float(sig[1::2].mean())},
        "adam_tied_sigma": {"steps": 200, "lr": 0.05, "classes": "even/odd observation index",
                            "J_first": hist2[0], "J_last": hist2[-1], "L": [L0, L2],
                            "inflation": [a0, a2], "sigma_true_0.05": float(sigc[0]),
                            "sigma_true_0.2": float(sigc[1])},
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--scale", type=int, default=1,
                    help="config 4: divide both horizontal grid sizes (and the profiles by "
                         "scale^2), keeping the local observation counts")
    ap.add_argument("--precision", default=None, choices=["fp64", "mixed", "both"],
                    help="default: both for config 2 (the headline: mixed timed against FP64), "
                         "mixed for config 4 (FP32 members), fp64 otherwise")
    ap.add_argument("--eig", default="ql", choices=["ql", "jacobi"],
                    help="eigensolver: tridiagonal QL (default) or one-sided Jacobi")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serial", action="store_true",
                    help="run every step with phase timing (zone chunks in order on one stream), "
                         "as the rocprofv3 kernel summaries are taken")
    ap.add_argument("--tune", action="store_true",
                    help="autotune the launch shapes (chunks in flight, zones per QL wave, HIP "
                         "graph) on this workload and store them in oak_amd/tuning.json")
    ap.add_argument("--calib", action="store_true",
                    help="also run a plain 1 GiB device copy (calibrates rocprof byte counters)")
    ap.add_argument("--calib-only", action="store_true",
                    help="run only the 1 GiB device copy and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    prec = args.precision or {2: "both", 4: "mixed"}.get(args.config, "fp64")
    import torch
    from oak_amd import Context, OAKG_FP64, OAKG_MIXED
    from oak_amd._lib import OAKG_EIG_JACOBI, OAKG_EIG_QL
    from oak_amd.errors import ensemble_errors
    device = torch.device("cuda:0")
    if args.calib_only:
        src = torch.ones(1 << 27, dtype=torch.float64, device=device)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print("# calib copy: 1073741824 bytes read, 1073741824 bytes written", file=sys.stderr)
        return
    if args.config == 3:
        run_config3(args, device)
        return
    t0 = time.time()
    DP, desc = build_problem(args.config, device, args.scale)
    P = DP.host
    g, H, Hshift, yo, rmse2 = device_inputs(DP, device)
    print("# generated %s in %.1f s" % (desc, time.time() - t0), file=sys.stderr, flush=True)
    n, N = DP.E.shape
    eig_name = (("k_tridiag_lower<256>+k_tridiag<128>+k_formq256+k_tql+k_tql_apply256+k_eigvec_epi (Householder on "
                 "register tiles, Q in blocked WY form on FP64 MFMA, implicit QL, rotations on "
                 "rows in registers, eigenvector epilogue; FP64)" if N > 128 else
                 "k_tridiag_lower<128>+k_tridiag_lower<64>+k_formq128+k_tql+k_tql_apply128 "
                 "(Householder on the lower triangle, Q in "
                 "blocked WY form on FP64 MFMA, implicit QL, eigenvector epilogue; FP64), with "
                 "the zones of p < 64 through k_tridiag_lower<64>+k_formq64+k_tql+k_tql_apply64+"
                 "k_eigvec_epi at 64 "
                 "and k_dual_lift" if N > 64 else
                 "k_tridiag_lower<64>+k_formq64+k_tql+k_tql_apply64+k_eigvec_epi (Householder on "
                 "the lower triangle + implicit QL, FP64; k_tridiag below N = 33)")
                if args.eig == "ql" else "k_eig (one-sided Jacobi, FP64)")
    eig_id = OAKG_EIG_QL if args.eig == "ql" else OAKG_EIG_JACOBI
    res = {}
    Ea64 = None
    if prec in ("fp64", "both"):
        ctx = Context(0, OAKG_FP64)
        ctx.set_eigensolver(eig_id)
        E = DP.E.double() if DP.E.dtype != torch.float64 else DP.E
        Ea64 = torch.empty((N, n), dtype=torch.float64, device=device).t()
        res["fp64"] = run_analysis(args, ctx, g, H, Hshift, yo, rmse2, E, Ea64, "fp64")
        counts = ctx.counts(P.nzones)
        ctx.close()
    if prec in ("mixed", "both"):
        ctx = Context(0, OAKG_MIXED)
        ctx.set_eigensolver(eig_id)
        if DP.E.dtype == torch.float32:  # FP32 members already (config 4)
            E32 = DP.E
        else:
            E32 = torch.empty((N, n), dtype=torch.float32, device=device).t()
            E32.copy_(DP.E)
        Ea32 = torch.empty((N, n), dtype=torch.float32, device=device).t()
        res["mixed"] = run_analysis(args, ctx, g, H, Hshift, yo, rmse2, E32, Ea32, "mixed")
        counts = ctx.counts(P.nzones)
        ctx.close()
        del E32
    head = "mixed" if "mixed" in res else "fp64"
    ms, phase, st_head = res[head]
    zs = np.diff(P.zone_ptr)
    F = flops_model(N, counts, zs, st_head.get("tuning", {}).get("dual", 0))
    out = {
        "metric": "analysed state elems/s",
        "value": n / (ms * 1e-3),
        "unit": "analysed state elems/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if head == "fp64" else "f32 state, f32 MFMA, f64 eigensolve",
        "data": "synthetic (seeded GRF ensembles and observations, random-init, no datasets)",
        "config": {"workload": "config%d: %s, %s, %s eigensolver" % (
                       args.config, desc, head, args.eig),
                   "n": int(n), "N": int(N), "m": int(P.m), "zones": int(P.nzones),
                   "mean_p": float(counts.mean()), "median_p": float(np.median(counts)),
                   "max_p": int(counts.max()),
                   "phase_ms": {k: round(v, 3) for k, v in phase.items()}},
        "roofline": roofline(F, phase, head == "mixed", eig_name, args.config, N),
        "roofline_phases": phase_rooflines(F, phase, head == "mixed", N),
        "launch": res[head][2].get("tuning"),
    }
    if "mixed" in res and "fp64" in res:
        ms64, ph64, st64 = res["fp64"]
        F64 = flops_model(N, counts, zs, st64.get("tuning", {}).get("dual", 0))
        out["fp64"] = {"value": n / (ms64 * 1e-3), "ms_per_step": ms64,
                       "phase_ms": {k: round(v, 3) for k, v in ph64.items()},
                       "roofline_phases": phase_rooflines(F64, ph64, False, N),
                       "launch": st64.get("tuning")}
        out["mixed_speedup_vs_fp64"] = ms64 / ms
        e = ensemble_errors(Ea32, Ea64, DP.E)
        out["mixed_err_vs_fp64"] = dict(e, bound=1e-4)
    if Ea64 is not None:
        out["report"] = run_report(args, H, Hshift, yo, rmse2, DP.E, Ea64)
    if args.calib:
        # a plain copy of known size: 1 GiB read + 1 GiB written
        src = torch.ones(1 << 27, dtype=torch.float64, device=device)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print("# calib copy: 1073741824 bytes read, 1073741824 bytes written", file=sys.stderr,
              flush=True)
    if not args.no_cpu:
        gpu_res = {}
        if Ea64 is not None:
            gpu_res["fp64"] = Ea64
        if "mixed" in res:
            gpu_res["mixed"] = Ea32
        out["cpu_baseline"], errs = cpu_baseline(DP, gpu_results=gpu_res)
        # the analyses against the FP64 oracle on the sampled zones (bounds:
        # 1e-10 for the FP64 path, 1e-4 for the mixed path)
        out["oracle_err"] = {k: dict(v, bound=1e-10 if k == "fp64" else 1e-4)
                             for k, v in errs.items()}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
