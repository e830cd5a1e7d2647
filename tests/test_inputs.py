"""Host-side input handling (CPU only): the reference's own JSON inputs and
LAMMPS data file (copied as fixtures under tests/golden/reference_inputs/),
JSON quirks, read_data semantics, special lists and the write/read round trip."""
import os

import numpy as np
import pytest

from rbmd_amd import make_water, read_data, run_config_from_json, write_data
from rbmd_amd.config import ConfigError, load_json, strip_json
from rbmd_amd.system import special_lists

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_inputs")


def test_strip_json_quirks():
    text = '﻿{ "a": 1, // comment\n "b": "x//y", /* block\n comment */ "c": [1, 2,], }'
    import json

    assert json.loads(strip_json(text)) == {"a": 1, "b": "x//y", "c": [1, 2]}


def test_reference_water_json():
    rc = run_config_from_json(load_json(os.path.join(GOLD, "H2O_rbmd.json")))
    assert rc.unit == "REAL" and not rc.inbuild
    assert rc.force_field == "CVFF"
    assert rc.neighbor_type == "VERLETLIST" and rc.cut_off == 10.0 and rc.r_core == 10.0
    assert rc.coulomb_type == "EWALD" and rc.rbe_p == 100 and rc.alpha == 0.0361 and rc.kmax == [3, 3, 3]
    assert rc.special_bonds == [0, 0, 0]
    assert rc.fix_shake == "null"  # commented out in the file (Configuration.h:98-110)
    assert rc.temperature == [298.0, 298.0, 50.0] and rc.temp_ctrl == "NOSE_HOOVER"
    assert rc.dt == 0.5 and rc.num_steps == 10000 and rc.thermo_interval == 1


def test_reference_salt_json_scalar_kmax_rejected():
    # the reference reads kmax as a vector and aborts on a scalar (Coulomb.cpp:40)
    with pytest.raises(ConfigError):
        run_config_from_json(load_json(os.path.join(GOLD, "LJ_salt_rbmd.json")))


def test_reference_water_data_file():
    s = read_data(os.path.join(GOLD, "H2O_212121_v.data"), "full", "REAL", special_bonds=[0, 0, 0])
    assert s.n_atoms == 300
    assert np.allclose(s.lo, -0.501) and np.allclose(s.hi, 20.509000778)
    assert len(s.bond_atoms) == 2 * 200 and len(s.angle_atoms) == 3 * 100
    assert set(np.unique(s.charge)) == {-0.82, 0.41}
    assert abs(s.charge.sum()) < 1e-12
    assert set(np.round(np.unique(s.mass), 4)) == {15.9994, 1.008}
    assert np.allclose(s.eps, [0.1553, 0.0]) and np.allclose(s.sigma, [3.166, 0.0])
    assert np.all(s.bond_k == 540.634) and np.all(s.bond_r0 == 1.0)
    assert np.all(s.angle_k == 50.0) and np.all(s.angle_theta0 == 109.47)
    assert s.velocities_read and np.abs(s.v).max() > 0
    # special lists with [0,0,0]: each atom lists its molecule partners with weight 0
    off, ids = s.special_offsets, s.special_ids
    for a in range(s.n_atoms):
        partners = set(ids[off[a]:off[a + 1]].tolist())
        same_mol = set(np.nonzero(s.mol_id == s.mol_id[a])[0].tolist()) - {a}
        assert partners == same_mol
    assert np.all(s.special_weights == 0.0)
    # angle triplets are H, O, H (the SHAKE worklet's order)
    ang = s.angle_atoms.reshape(-1, 3)
    assert np.all(s.type[ang[:, 1]] == 0) and np.all(s.type[ang[:, 0]] == 1) and np.all(s.type[ang[:, 2]] == 1)


def test_special_lists_chain():
    # 0-1-2-3 chain: 1-2, 1-3 and 1-4 partners get weights w0, w1, w2
    off, ids, w = special_lists(5, np.array([0, 1, 1, 2, 2, 3], np.int32), [5, 6, 7])
    got = {a: list(zip(ids[off[a]:off[a + 1]].tolist(), w[off[a]:off[a + 1]].tolist())) for a in range(5)}
    assert got[0] == [(1, 5.0), (2, 6.0), (3, 7.0)]
    assert got[4] == [(4, 1.0)]  # an atom without bonds lists itself with weight 1
    assert (1, 5.0) in got[2] and (3, 5.0) in got[2] and (0, 6.0) in got[2]


def test_write_read_round_trip(tmp_path):
    s = make_water(n_mol=50, seed=3)
    p = tmp_path / "w.data"
    write_data(str(p), s)
    r = read_data(str(p), "full", "REAL", special_bonds=[0, 0, 0])
    assert np.array_equal(r.x, s.x) and np.array_equal(r.v, s.v)
    assert np.array_equal(r.type, s.type) and np.array_equal(r.charge, s.charge)
    assert np.array_equal(r.bond_atoms, s.bond_atoms) and np.array_equal(r.angle_atoms, s.angle_atoms)
    assert np.array_equal(r.special_ids, s.special_ids)


def test_water_special_lists_fast_path_matches_generic():
    from rbmd_amd.system import water_special_lists

    n_mol = 7
    base = 3 * np.arange(n_mol)
    bonds = np.stack([base, base + 1, base, base + 2], 1).reshape(-1).astype(np.int32)
    for sb in ([0, 0, 0], [0, 1, 1], [2, 3, 4]):
        a = special_lists(3 * n_mol, bonds, sb)
        b = water_special_lists(n_mol, sb)
        for u, v in zip(a, b):
            assert np.array_equal(u, v)


def test_rigid_water_velocities():
    s = make_water(n_mol=200, seed=2)
    L = s.box[0]
    x = s.x.reshape(-1, 3, 3)
    v = s.v.reshape(-1, 3, 3)
    d = x[:, 1:] - x[:, :1]
    d -= L * np.round(d / L)
    dv = v[:, 1:] - v[:, :1]
    # rigid motion: no stretching along the O-H bonds
    assert np.abs((d * dv).sum(2)).max() < 1e-12 * max(1.0, np.abs(v).max())
