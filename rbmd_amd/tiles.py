"""Tile choices measured by rbmd_amd/tune.py (rbmd_amd/tiles.json).

The table holds, per force field and precision, the wavefronts that split
the k vectors of the k-space force kernel and the k vectors per lane of the
structure-factor kernel (one choice per kind and precision, so every system
of a pack shares them) and the pair kernel's lanes per atom by atom count (a
per-system runtime value).
Without a table the library's defaults apply.
"""
import json
import os

TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tiles.json")
_cache = None


def load():
    global _cache
    if _cache is None:
        _cache = {}
        if os.path.exists(TABLE):
            with open(TABLE) as f:
                _cache = json.load(f)
    return _cache


def choose(force_field: str, precision: str, n_atoms: int) -> dict:
    """{"pair_lanes", "kforce_lanes", "rho_kblock"} for one system (0: default)"""
    entry = load().get("kinds", {}).get(force_field, {}).get(precision)
    if not entry:
        return {}
    out = {"kforce_lanes": int(entry.get("kforce_lanes", 0)), "rho_kblock": int(entry.get("rho_kblock", 0))}
    lanes = 0
    for max_atoms, value in entry.get("pair_lanes", []):
        lanes = int(value)
        if n_atoms <= max_atoms:
            break
    out["pair_lanes"] = lanes
    return out
