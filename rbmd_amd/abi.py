"""ctypes mirror of the plain-C types in include/rbmd_mi355x.h.

These are data layouts only (no calls).  The field order must match the C
header exactly; tests/test_abi.py checks the sizes against the compiled
library's own sizeof values.
"""
import ctypes as C

import numpy as np

RBMD_OK = 0
RBMD_ERR_ARG = -1
RBMD_ERR_OOM = -2
RBMD_ERR_HIP = -3
RBMD_ERR_SHAKE = -4
RBMD_ERR_NONFINITE = -5
RBMD_ERR_STATE = -6
RBMD_ERR_CAPACITY = -7
RBMD_ERR_UNSUPPORTED = -8
RBMD_ERR_NODEVICE = -9

UNIT = {"LJ": 0, "REAL": 1, "METAL": 2}
FORCE_FIELD = {"LJ/CUT": 0, "LJ/CUT/COUL/LONG": 1, "CVFF": 2}
NEIGHBOR = {"VERLETLIST": 0, "RBL": 1}
COULOMB = {"NONE": 0, "RBE": 1, "EWALD": 2}
SHAKE = {"null": 0, "true": 1, "false": 2}
TEMP_CTRL = {"NOSE_HOOVER": 0, "BERENDSEN": 1, "TEMP_RESCALE": 2, "NONE": 3}
PRECISION = {"fp64": 0, "mixed": 1}
ERFC = {"analytic": 0, "table": 1}
PAIR_PATH = {"lists": 0, "cells": 1}

_i32p = C.POINTER(C.c_int32)
_f64p = C.POINTER(C.c_double)


class SysDesc(C.Structure):
    _fields_ = [
        ("n_atoms", C.c_int32),
        ("n_types", C.c_int32),
        ("lo", C.c_double * 3),
        ("hi", C.c_double * 3),
        ("type", _i32p),
        ("charge", _f64p),
        ("mass", _f64p),
        ("mol_id", _i32p),
        ("eps", _f64p),
        ("sigma", _f64p),
        ("special_offsets", _i32p),
        ("special_ids", _i32p),
        ("special_weights", _f64p),
        ("n_bonds", C.c_int32),
        ("bond_atoms", _i32p),
        ("bond_k", _f64p),
        ("bond_r0", _f64p),
        ("n_angles", C.c_int32),
        ("angle_atoms", _i32p),
        ("angle_k", _f64p),
        ("angle_theta0", _f64p),
    ]


class Params(C.Structure):
    _fields_ = [
        ("unit", C.c_int32),
        ("force_field", C.c_int32),
        ("neighbor_type", C.c_int32),
        ("coulomb_type", C.c_int32),
        ("cut_off", C.c_double),
        ("r_core", C.c_double),
        ("neighbor_sample_num", C.c_int32),
        ("rbe_p", C.c_int32),
        ("alpha", C.c_double),
        ("kmax", C.c_int32 * 3),
        ("fix_shake", C.c_int32),
        ("special_bonds", C.c_int32),
        ("inbuild", C.c_int32),
        ("temp_ctrl", C.c_int32),
        ("temperature", C.c_double * 3),
        ("dt", C.c_double),
        ("num_steps", C.c_int64),
        ("precision", C.c_int32),
        ("erfc_mode", C.c_int32),
        ("seed", C.c_uint64),
        ("energy_every", C.c_int32),
        ("rbl_rho", C.c_double),
        ("pair_path", C.c_int32),
        ("pair_lanes", C.c_int32),
        ("kforce_lanes", C.c_int32),
        ("rho_kblock", C.c_int32),
        ("shake_tol", C.c_double),
        ("skin", C.c_double),
    ]


class Thermo(C.Structure):
    _fields_ = [
        ("step", C.c_int64),
        ("temperature", C.c_double),
        ("pressure", C.c_double),
        ("ke_sum", C.c_double),
        ("e_lj", C.c_double),
        ("e_coul", C.c_double),
        ("e_long", C.c_double),
        ("e_bond", C.c_double),
        ("e_angle", C.c_double),
        ("virial", C.c_double * 6),
        ("nhc_eta", C.c_double * 3),
        ("nhc_eta_dot", C.c_double * 3),
        ("energy_step", C.c_int64),
        ("shake_max_iter", C.c_int32),
        ("status_flags", C.c_int32),
        ("misfit_atoms", C.c_int32),
        ("list_builds", C.c_int32),
    ]

    def as_dict(self):
        out = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            if isinstance(v, C.Array):
                v = list(v)
            out[name] = v
        return out


def ptr(a, ctype):
    """Pointer to a contiguous numpy array (or NULL for None)."""
    if a is None:
        return C.cast(None, C.POINTER(ctype))
    return a.ctypes.data_as(C.POINTER(ctype))


class DescHolder:
    """Builds a SysDesc from a System and keeps the numpy buffers alive."""

    def __init__(self, system):
        s = system
        n = s.n_atoms
        self._keep = []

        def arr(a, dtype):
            if a is None:
                return None
            b = np.ascontiguousarray(a, dtype=dtype)
            self._keep.append(b)
            return b

        d = SysDesc()
        d.n_atoms = n
        d.n_types = len(s.eps)
        for k in range(3):
            d.lo[k] = float(s.lo[k])
            d.hi[k] = float(s.hi[k])
        d.type = ptr(arr(s.type, np.int32), C.c_int32)
        d.charge = ptr(arr(s.charge, np.float64), C.c_double)
        d.mass = ptr(arr(s.mass, np.float64), C.c_double)
        d.mol_id = ptr(arr(s.mol_id, np.int32), C.c_int32)
        d.eps = ptr(arr(s.eps, np.float64), C.c_double)
        d.sigma = ptr(arr(s.sigma, np.float64), C.c_double)
        if s.special_offsets is not None:
            d.special_offsets = ptr(arr(s.special_offsets, np.int32), C.c_int32)
            d.special_ids = ptr(arr(s.special_ids, np.int32), C.c_int32)
            d.special_weights = ptr(arr(s.special_weights, np.float64), C.c_double)
        else:
            d.special_offsets = ptr(None, C.c_int32)
            d.special_ids = ptr(None, C.c_int32)
            d.special_weights = ptr(None, C.c_double)
        nb = 0 if s.bond_atoms is None else len(s.bond_atoms) // 2
        d.n_bonds = nb
        d.bond_atoms = ptr(arr(s.bond_atoms, np.int32) if nb else None, C.c_int32)
        d.bond_k = ptr(arr(s.bond_k, np.float64) if nb else None, C.c_double)
        d.bond_r0 = ptr(arr(s.bond_r0, np.float64) if nb else None, C.c_double)
        na = 0 if s.angle_atoms is None else len(s.angle_atoms) // 3
        d.n_angles = na
        d.angle_atoms = ptr(arr(s.angle_atoms, np.int32) if na else None, C.c_int32)
        d.angle_k = ptr(arr(s.angle_k, np.float64) if na else None, C.c_double)
        d.angle_theta0 = ptr(arr(s.angle_theta0, np.float64) if na else None, C.c_double)
        self.desc = d


def make_params(run, n_atoms=None):
    """Params from a RunConfig (rbmd_amd.config).  Tile choices: the
    RunConfig's explicit values, else (run.tiles == "tuned" and n_atoms given)
    the tuned table rbmd_amd/tiles.json, else the library's defaults (0)."""
    p = Params()
    p.unit = UNIT[run.unit]
    p.force_field = FORCE_FIELD[run.force_field]
    p.neighbor_type = NEIGHBOR[run.neighbor_type]
    p.coulomb_type = COULOMB[run.coulomb_type]
    p.cut_off = run.cut_off
    p.r_core = run.r_core if run.neighbor_type == "RBL" else run.cut_off
    p.neighbor_sample_num = int(run.neighbor_sample_num)
    p.rbe_p = int(run.rbe_p)
    p.alpha = run.alpha
    for k in range(3):
        p.kmax[k] = int(run.kmax[k])
    p.fix_shake = SHAKE[run.fix_shake]
    p.special_bonds = 1 if run.special_bonds else 0
    p.inbuild = 1 if run.inbuild else 0
    p.temp_ctrl = TEMP_CTRL[run.temp_ctrl]
    for k in range(3):
        p.temperature[k] = run.temperature[k]
    p.dt = run.dt
    p.num_steps = int(run.num_steps)
    p.precision = PRECISION[run.precision]
    p.erfc_mode = ERFC[run.erfc_mode]
    p.seed = int(run.seed) & 0xFFFFFFFFFFFFFFFF
    p.energy_every = int(run.energy_every)
    p.rbl_rho = float(run.rbl_rho)
    p.pair_path = PAIR_PATH[run.pair_path]
    tiles = {"pair_lanes": 0, "kforce_lanes": 0, "rho_kblock": 0}
    if getattr(run, "tiles", "tuned") == "tuned" and n_atoms is not None:
        from .tiles import choose

        tiles.update(choose(run.force_field, run.precision, int(n_atoms)))
    for k in tiles:
        v = int(getattr(run, k, 0) or 0)
        setattr(p, k, v if v else int(tiles[k]))
    p.shake_tol = float(getattr(run, "shake_tol", 0.0) or 0.0)
    p.skin = float(getattr(run, "skin", 0.0) or 0.0)
    return p
