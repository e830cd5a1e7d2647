"""The `rbmd -j` driver (rbmd_amd.run) on the CPU oracle: the executioner's
step count and output schedule, and the formats of energy.rbmd,
temperature.rbmd and rbmd.trj (ThermoOutput.cpp:223-291, TempOutput.cpp:60-90,
TrajectoryOutput.cpp:44-80).  The GPU run is compared with this one in
tests/test_gpu_run_driver.py."""
import os

import numpy as np

import oracle
from rbmd_amd.config import load_json, run_config_from_json
from rbmd_amd.run import THERMO_HEADER, THERMO_HEADER_MOL, cfmt, load_system, run_md

HERE = os.path.dirname(os.path.abspath(__file__))
REF = os.path.join(HERE, "golden", "reference_inputs", "H2O_rbmd.json")
INBUILD = os.path.join(HERE, "golden", "run_inputs", "lj_inbuild.json")


def drive(path, out, steps=None, sim_factory=None):
    doc = load_json(path)
    run = run_config_from_json(doc)
    run.energy_every = 0
    system = load_system(doc, os.path.dirname(path), run)
    n = run.num_steps if steps is None else steps
    sim = (sim_factory or oracle.OracleSystem)(system, run)
    last = run_md(sim, system, run, doc, str(out), n)
    return system, run, last, sim


def rows(path):
    with open(path) as f:
        lines = f.read().splitlines()
    return lines[0], [ln.split() for ln in lines[1:]]


def test_cfmt_matches_ostream_default():
    assert cfmt(3) == "3"
    assert cfmt(0.1) == "0.1"
    assert cfmt(-1234567.0) == "-1.23457e+06"
    assert cfmt(1e-5) == "1e-05"
    assert cfmt(298.0) == "298"


def test_water_reference_input(tmp_path):
    system, run, last, sim = drive(REF, tmp_path, steps=6)
    head, body = rows(tmp_path / "energy.rbmd")
    assert head == THERMO_HEADER + THERMO_HEADER_MOL
    # interval 1: rows for k = 0..6, 15 columns (plus the trailing space the reference writes)
    assert [int(r[0]) for r in body] == list(range(7))
    assert all(len(r) == 15 for r in body)
    assert float(body[0][5]) == 0.0  # Residual at step 0
    th, tb = rows(tmp_path / "temperature.rbmd")
    assert th == "Step Temperature Pressure" and [int(r[0]) for r in tb] == list(range(7))
    # interval 100 > 6 steps: no trajectory frame
    assert os.path.getsize(tmp_path / "rbmd.trj") == 0
    # the run took num_steps + 1 = 7 steps; the last row is the state after 6
    assert sim.thermo()["step"] == 7
    assert abs(float(body[-1][2]) - last["lj"]) <= 1e-5 * abs(last["lj"]) + 1e-12
    # PE = non-bonded + bond + angle; total = PE + KE
    pe, ke, te = float(body[-1][7]), float(body[-1][6]), float(body[-1][8])
    assert abs(te - (pe + ke)) <= 1e-5 * abs(te)


def test_inbuild_lj(tmp_path):
    system, run, last, sim = drive(INBUILD, tmp_path)
    assert system.n_atoms == 216 and np.allclose(system.x[1] - system.x[0], [1.2, 0, 0])
    head, body = rows(tmp_path / "energy.rbmd")
    assert head == THERMO_HEADER  # in-build: no molecular columns
    assert [int(r[0]) for r in body] == [0, 5, 10, 15, 20]
    assert all(len(r) == 10 for r in body)
    assert all(float(r[3]) == 0.0 and float(r[4]) == 0.0 for r in body)  # LJ/CUT: no Coulomb columns
    with open(tmp_path / "rbmd.trj") as f:
        lines = f.read().splitlines()
    frames = [i for i, ln in enumerate(lines) if ln == "ITEM: TIMESTEP"]
    assert [int(lines[i + 1]) for i in frames] == [10, 20]
    assert lines[frames[0] + 3] == "216" and lines[frames[0] + 4] == "ITEM: BOX BOUNDS pp pp pp"
    assert lines[frames[0] + 5] == "0 7.2"
    first = lines[frames[0] + 9].split()
    assert first[:2] == ["1", "1"] and len(first) == 5
    assert sim.thermo()["step"] == 21
