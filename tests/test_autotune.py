"""The launch-shape table of oak_amd/autotune.py (no GPU): the candidate
space per ensemble size, the table entries and their defaults."""
import json

from oak_amd import autotune


def test_space_leaves_out_the_wide_kernel_knobs_at_64():
    s64 = autotune.space_for(64)
    s128 = autotune.space_for(100)
    assert set(s128) == set(autotune.CANDIDATES)
    assert not set(autotune.WIDE_ONLY) & set(s64)
    assert set(s64) | set(autotune.WIDE_ONLY) == set(s128)
    # N > 128 has its own kernels: the N = 128 options do nothing there
    assert set(autotune.space_for(256)) == set(s64)
    # every default is a candidate, and every candidate is a value the
    # library takes (oakg_set_tuning ranges)
    ok = {"slots": range(1, 5), "ql_zones": (4, 8, 16, 32), "graph": (0, 1), "order": (0, 1),
          "tri_split": (0, 1, 2, 3, 4), "dual": (0, 16, 32, 64), "persist": (0, 1, 2)}
    for k, vals in autotune.CANDIDATES.items():
        assert autotune.DEFAULT[k] in vals or k == "ql_zones"
        assert all(v in ok[k] for v in vals), k


def test_table_round_trip(tmp_path):
    path = str(tmp_path / "t.json")
    assert autotune.load_table(path) == {}
    best = dict(autotune.DEFAULT, persist=2, dual=64)
    times = {json.dumps(best, sort_keys=True): 70.1234, json.dumps(autotune.DEFAULT): 72.0}
    autotune.save_entry(128, 1, best, times, "config2", path=path)
    tab = autotune.load_table(path)
    ent = tab[autotune.key(128, 1)]
    assert ent["settings"] == best and ent["workload"] == "config2"
    assert list(ent["ms"].values()) == [70.123, 72.0]  # fastest first
    assert autotune.key(64, 0) == "N64_fp64"


def test_shipped_table_entries_are_valid():
    tab = autotune.load_table()
    assert tab, "oak_amd/tuning.json is missing"
    for k, ent in tab.items():
        N = int(k.split("_")[0][1:])
        for name, v in ent["settings"].items():
            assert name in autotune.CANDIDATES, (k, name)
            if N <= 64:
                assert name not in autotune.WIDE_ONLY or v == autotune.DEFAULT[name], (k, name)
