"""Run configuration: the `rbmd -j <file>.json` contract.

The JSON blocks and keys are the reference's
(framework/application/MDApplication.cpp:104-244, hyperparameters/*.cpp):
``init_configuration{read_data|inbuild}``, ``hyper_parameters{force_field,
neighbor, coulomb, extend}``, ``execution{...}``, ``outputs{...}``.
The reference parses with jsoncpp, which accepts // and /* */ comments,
trailing commas and a UTF-8 BOM (e.g. res/json_files/LJ_salt/rbmd.json:1,29);
``load_json`` accepts the same.  Options that only this build has live in a
separate ``"mi355x"`` block so reference inputs parse unchanged.
"""
from __future__ import annotations

import dataclasses
import json
import os
import re
from typing import List


@dataclasses.dataclass
class RunConfig:
    unit: str = "LJ"
    force_field: str = "LJ/CUT/COUL/LONG"
    neighbor_type: str = "VERLETLIST"
    cut_off: float = 4.0
    r_core: float = 4.0
    neighbor_sample_num: int = 100
    coulomb_type: str = "RBE"
    rbe_p: int = 100
    alpha: float = 0.64
    kmax: List[int] = dataclasses.field(default_factory=lambda: [4, 4, 4])
    fix_shake: str = "null"
    special_bonds: list = None
    inbuild: bool = False
    temp_ctrl: str = "NOSE_HOOVER"
    temperature: List[float] = dataclasses.field(default_factory=lambda: [1.5, 1.5, 0.1])
    dt: float = 0.001
    num_steps: int = 1000
    precision: str = "fp64"
    erfc_mode: str = "analytic"
    seed: int = 12345
    energy_every: int = 0
    rbl_rho: float = 0.0
    pair_path: str = "lists"
    tiles: str = "tuned"  # "tuned": rbmd_amd/tiles.json by kind, precision and size; "default": library defaults
    pair_lanes: int = 0  # explicit tile choices (0: from `tiles`)
    kforce_lanes: int = 0
    rho_kblock: int = 0
    thermo_interval: int = 0
    shake_tol: float = 0.0  # 0: the reference's 1e-5 on the multiplier change
    skin: float = 0.0  # Verlet-list skin of MD steps (0: a list per step, as the reference)


def strip_json(text: str) -> str:
    """Remove BOM, // and /* */ comments (outside strings) and trailing commas."""
    if text.startswith("﻿"):
        text = text[1:]
    out = []
    i, n = 0, len(text)
    in_str = False
    while i < n:
        c = text[i]
        if in_str:
            out.append(c)
            if c == "\\" and i + 1 < n:
                out.append(text[i + 1])
                i += 2
                continue
            if c == '"':
                in_str = False
            i += 1
            continue
        if c == '"':
            in_str = True
            out.append(c)
            i += 1
        elif text.startswith("//", i):
            j = text.find("\n", i)
            i = n if j < 0 else j
        elif text.startswith("/*", i):
            j = text.find("*/", i + 2)
            i = n if j < 0 else j + 2
        else:
            out.append(c)
            i += 1
    s = "".join(out)
    return re.sub(r",(\s*[}\]])", r"\1", s)


def load_json(path: str) -> dict:
    with open(path, "r", encoding="utf-8") as f:
        return json.loads(strip_json(f.read()))


class ConfigError(ValueError):
    pass


def _req(node: dict, key: str, where: str):
    if key not in node:
        # the reference prints and exits with status 0 (Logging.h:44-51); here it raises
        raise ConfigError(f"missing key '{key}' in {where}")
    return node[key]


def run_config_from_json(doc: dict) -> RunConfig:
    rc = RunConfig()
    init = _req(doc, "init_configuration", "root")
    if "read_data" in init:
        rd = init["read_data"]
        rc.unit = _req(rd, "unit", "read_data")
        rc.inbuild = False
    elif "inbuild" in init:
        rc.unit = _req(init["inbuild"], "unit", "inbuild")
        rc.inbuild = True
    hp = _req(doc, "hyper_parameters", "root")
    ff = _req(hp, "force_field", "hyper_parameters")
    rc.force_field = _req(ff, "type", "force_field")
    nb = _req(hp, "neighbor", "hyper_parameters")
    rc.neighbor_type = _req(nb, "type", "neighbor")
    rc.cut_off = float(_req(nb, "cut_off", "neighbor"))
    if rc.neighbor_type == "RBL":  # Neighbor.cpp:40-48
        rc.r_core = float(_req(nb, "r_core", "neighbor"))
        rc.neighbor_sample_num = int(float(_req(nb, "neighbor_sample_num", "neighbor")))
    else:
        rc.r_core = rc.cut_off
    if "coulomb" in hp:
        cb = hp["coulomb"]
        rc.coulomb_type = _req(cb, "type", "coulomb")
        rc.rbe_p = int(_req(cb, "coulomb_sample_num", "coulomb"))
        rc.alpha = float(_req(cb, "alpha", "coulomb"))
        km = _req(cb, "kmax", "coulomb")
        if not isinstance(km, list):  # Coulomb.cpp:40 reads a vector (hazard Q22)
            raise ConfigError("coulomb.kmax must be an array of 3 integers")
        rc.kmax = [int(k) for k in km]
    else:
        rc.coulomb_type = "NONE"
    ext = hp.get("extend", {}) or {}
    fs = ext.get("fix_shake", None)
    rc.fix_shake = "null" if fs is None else ("true" if fs in (True, "true") else "false")
    sb = ext.get("special_bonds", None)
    rc.special_bonds = [int(w) for w in sb] if sb else None
    ex = _req(doc, "execution", "root")
    rc.temperature = [float(t) for t in _req(ex, "temperature", "execution")]
    rc.temp_ctrl = _req(ex, "temp_ctrl_type", "execution")
    rc.dt = float(_req(ex, "timestep", "execution"))
    rc.num_steps = int(float(_req(ex, "num_steps", "execution")))
    if ex.get("ensemble", "NVT") == "NVE":
        rc.temp_ctrl = "NONE"
    outs = doc.get("outputs", {}) or {}
    if "thermo_out" in outs:
        rc.thermo_interval = int(outs["thermo_out"].get("interval", 1))
    ext2 = doc.get("mi355x", {}) or {}
    rc.precision = ext2.get("precision", rc.precision)
    rc.erfc_mode = ext2.get("erfc", "table" if rc.neighbor_type == "RBL" else "analytic")
    rc.seed = int(ext2.get("seed", rc.seed))
    rc.energy_every = int(ext2.get("energy_every", rc.thermo_interval))
    rc.rbl_rho = float(ext2.get("rbl_rho", 0.0))
    rc.pair_path = str(ext2.get("pair_path", "lists"))
    rc.tiles = str(ext2.get("tiles", "tuned"))
    for key in ("pair_lanes", "kforce_lanes", "rho_kblock"):
        setattr(rc, key, int(ext2.get(key, 0)))
    rc.shake_tol = float(ext2.get("shake_tol", 0.0))
    rc.skin = float(ext2.get("skin", 0.0))
    return rc


def salt_config(**kw) -> RunConfig:
    """Charged-LJ salt workload: LJ units, erfc cut 4 sigma (the reference has a single
    cutoff for LJ and erfc), beta = 0.8 -> alpha = beta^2 = 0.64, RBE P = 100,
    Nose-Hoover NVT T* = 1.5 with damping 0.1 tau, dt = 0.001, 1000 steps."""
    rc = RunConfig(unit="LJ", force_field="LJ/CUT/COUL/LONG", neighbor_type="VERLETLIST", cut_off=4.0,
                   r_core=4.0, coulomb_type="RBE", rbe_p=100, alpha=0.64, kmax=[6, 6, 6], fix_shake="null",
                   special_bonds=None, temp_ctrl="NOSE_HOOVER", temperature=[1.5, 1.5, 0.1], dt=0.001,
                   num_steps=1000)
    for k, v in kw.items():
        setattr(rc, k, v)
    return rc


def water_config(**kw) -> RunConfig:
    """Rigid 3-site water workload: REAL units, cut 10 A, beta = 0.3 -> alpha = 0.09,
    RBE P = 200, SHAKE, special bonds [0,0,0], Nose-Hoover 300 K with damping 100 fs,
    dt = 2 fs, 5000 steps."""
    rc = RunConfig(unit="REAL", force_field="CVFF", neighbor_type="VERLETLIST", cut_off=10.0, r_core=10.0,
                   coulomb_type="RBE", rbe_p=200, alpha=0.09, kmax=[8, 8, 8], fix_shake="true",
                   special_bonds=[0, 0, 0], temp_ctrl="NOSE_HOOVER", temperature=[300.0, 300.0, 100.0], dt=2.0,
                   num_steps=5000)
    for k, v in kw.items():
        setattr(rc, k, v)
    return rc
