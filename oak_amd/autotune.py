"""Autotuner of the zone loop's launch shapes.

The knobs (oakg_set_tuning) change how the zones are scheduled:
  slots      zone chunks in flight on as many streams,
  ql_zones   zones per wave of the QL kernel,
  graph      the zone loop replayed from a HIP graph,
  order      zones in order of decreasing observation count (the heaviest
             first, similar zones side by side in a QL wave),
  tri_split  the N > 64 Householder reduction as two kernels (the second
             with fewer registers and more zones per CU), three (the
             middle one on 6 x 6 register tiles), or two with the first on
             the lower triangle only (3), or both on it (4),
  dual       the bucket bound below which a zone is solved in its p x p
             observation space,
  persist    the Gram kernel as a persistent grid taking zones from a work
             counter (2: a zone per workgroup).
tri_split, dual and persist act at 64 < N <= 128 only (``space_for``
leaves them out elsewhere).  All but dual and tri_split = 2, 3, 4 give the same bits; those
give the same analysis to rounding.  ``tune`` times every candidate on the caller's own workload
(device-resident inputs, the same call repeated), keeps the fastest, and
returns a table entry; ``apply`` sets a context from a table (a JSON file
keyed by ensemble size and precision, with the workload it was tuned on)."""
import itertools
import json
import os
import time

CANDIDATES = {"slots": (2, 3, 4), "ql_zones": (4, 8, 16), "graph": (0, 1), "order": (0, 1),
              "tri_split": (0, 1, 2, 3, 4), "dual": (0, 32, 64), "persist": (0, 2)}
DEFAULT = {"slots": 3, "ql_zones": 16, "graph": 0, "order": 0, "tri_split": 0, "dual": 0,
           "persist": 0}
WIDE_ONLY = ("tri_split", "dual", "persist")  # knobs of the N > 64 kernels

TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")


def key(N, precision):
    return "N%d_%s" % (int(N), "mixed" if int(precision) == 1 else "fp64")


def space_for(N):
    """The candidates that change anything at ensemble size N."""
    if 64 < int(N) <= 128:
        return dict(CANDIDATES)
    return {k: v for k, v in CANDIDATES.items() if k not in WIDE_ONLY}


def tune(run, ctx, repeats=2, space=None, log=None):
    """run(): one call of the analysis with ctx (synchronised by the caller's
    function).  Every point of the space is run once untimed, then timed
    over ``repeats`` calls; returns (best settings, {settings: ms})."""
    import torch
    space = space or CANDIDATES
    names = list(space)
    times = {}
    for vals in itertools.product(*(space[n] for n in names)):
        s = dict(zip(names, vals))
        ctx.set_tuning(**s)
        run()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(repeats):
            run()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / repeats
        times[json.dumps(s, sort_keys=True)] = ms
        if log:
            log("# autotune %s: %.3f ms" % (s, ms))
    best = json.loads(min(times, key=times.get))
    ctx.set_tuning(**best)
    return best, times


def load_table(path=TABLE):
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def apply(ctx, N, precision, path=TABLE):
    """Sets ctx from the table entry of (N, precision); the defaults when
    the table has none.  Returns the settings applied."""
    ent = load_table(path).get(key(N, precision))
    s = dict(DEFAULT)
    if ent:
        s.update({k: ent["settings"][k] for k in DEFAULT if k in ent["settings"]})
    ctx.set_tuning(**s)
    return s


def save_entry(N, precision, settings, times, workload, path=TABLE):
    """Stores the entry of (N, precision); its timings fastest first."""
    tab = load_table(path)
    tab[key(N, precision)] = {"settings": dict(sorted(settings.items())), "workload": workload,
                              "ms": {k: round(v, 3) for k, v in sorted(times.items(),
                                                                       key=lambda kv: kv[1])}}
    with open(path, "w") as f:
        json.dump(dict(sorted(tab.items())), f, indent=1)
