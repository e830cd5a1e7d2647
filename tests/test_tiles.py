"""Tile choices (CPU): the tuner's choice rule on synthetic timings, the
table lookup by size, and the defaults without a table."""
from rbmd_amd import tiles, tune
from rbmd_amd.abi import make_params
from rbmd_amd.config import salt_config


def test_choose_rule():
    per_size = [
        (1000, {"pair_lanes": [30, 20, 10], "kforce_lanes": [20, 10, 12], "rho_kblock": [9, 12, 20]}),
        (100000, {"pair_lanes": [100, 90, 120], "kforce_lanes": [40, 39, 38], "rho_kblock": [30, 26, 40]}),
    ]
    e = tune.choose(per_size)
    assert e["pair_lanes"] == [[10000, 16], [2 ** 31 - 1, 8]]
    # kforce: relative sums 2+1.05 / 1+1.03 / 1.2+1 -> option 1 (4 lanes)
    assert e["kforce_lanes"] == 4
    # rho: 1+1.15 / 1.33+1 / 2.2+1.5 -> option 0 (1 k vector per lane)
    assert e["rho_kblock"] == 1


def test_choose_k_tiles_by_step_time():
    """where step times are measured they decide the k-space tiles, not the kernel times"""
    per_size = [
        (1000, {"pair_lanes": [30, 20, 10], "kforce_lanes": [20, 10, 12], "rho_kblock": [9, 12, 20],
                "step_kforce_lanes": [50, 52, 51], "step_rho_kblock": [60, 59, 58]}),
        (100000, {"pair_lanes": [100, 90, 120], "kforce_lanes": [40, 39, 38], "rho_kblock": [30, 26, 40],
                  "step_kforce_lanes": [300, 310, 320], "step_rho_kblock": [300, 290, 280]}),
    ]
    e = tune.choose(per_size)
    assert e["kforce_lanes"] == 2  # by kernel time it would be 4
    assert e["rho_kblock"] == 4  # by kernel time it would be 1
    assert e["pair_lanes"] == [[10000, 16], [2 ** 31 - 1, 8]]


def test_lookup_by_size(monkeypatch):
    table = {"kinds": {"CVFF": {"mixed": {"kforce_lanes": 8, "rho_kblock": 2,
                                          "pair_lanes": [[5000, 16], [50000, 8], [2 ** 31 - 1, 4]]}}}}
    monkeypatch.setattr(tiles, "_cache", table)
    assert tiles.choose("CVFF", "mixed", 3000) == {"kforce_lanes": 8, "rho_kblock": 2, "pair_lanes": 16}
    assert tiles.choose("CVFF", "mixed", 5001)["pair_lanes"] == 8
    assert tiles.choose("CVFF", "mixed", 10 ** 6)["pair_lanes"] == 4
    assert tiles.choose("CVFF", "fp64", 3000) == {}


def test_params_explicit_tuned_default(monkeypatch):
    table = {"kinds": {"LJ/CUT/COUL/LONG": {"fp64": {"kforce_lanes": 2, "rho_kblock": 4,
                                                     "pair_lanes": [[2 ** 31 - 1, 8]]}}}}
    monkeypatch.setattr(tiles, "_cache", table)
    p = make_params(salt_config(), 4000)
    assert (p.pair_lanes, p.kforce_lanes, p.rho_kblock) == (8, 2, 4)
    p = make_params(salt_config(pair_lanes=4), 4000)
    assert (p.pair_lanes, p.kforce_lanes, p.rho_kblock) == (4, 2, 4)
    p = make_params(salt_config(tiles="default"), 4000)
    assert (p.pair_lanes, p.kforce_lanes, p.rho_kblock) == (0, 0, 0)
    p = make_params(salt_config())  # no atom count: library defaults
    assert (p.pair_lanes, p.kforce_lanes, p.rho_kblock) == (0, 0, 0)


def test_committed_table_is_consistent():
    t = tiles.load()
    for ff, by_prec in t.get("kinds", {}).items():
        for prec, e in by_prec.items():
            assert e["kforce_lanes"] in (2, 4, 8) and e["rho_kblock"] in (1, 2, 4)
            edges = [edge for edge, _ in e["pair_lanes"]]
            assert edges == sorted(edges) and all(v in (4, 8, 16) for _, v in e["pair_lanes"])
