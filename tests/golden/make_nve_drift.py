"""Write the oracle's NVE energy-drift record for the salt workload.

NVE in the reference is a stub (framework/run/ExecutionNVE.cpp:36-41); its
parity form is the NVT step with the thermostat decoupled, plain velocity
Verlet (temp_ctrl NONE here, ExecutionNVT.cpp:1112-1116).  The workload is
the charged-LJ salt of 4,000 ions (L = 20, rc = 4, alpha = 0.64, dt = 0.001)
run for 10,000 steps twice: with the Ewald k-space sum over the |m| <= 6
grid (a deterministic force; the energy is the grid energy of the same grid)
and with RBE at P = 100 (seeded k sets, the same stream the product draws).
Every 100 steps the total energy per atom (kinetic + LJ + real-space
Coulomb + k-space + self) is recorded.  The drift is the least-squares slope
of that series in energy per atom per time unit, and the spread is
max |E - E0| / |E0|.

tests/test_gpu_nve_drift.py runs the same steps on the GPU and compares.
This takes about 25 minutes of CPU (8 threads):
    python tests/golden/make_nve_drift.py
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from rbmd_amd import make_salt, salt_config  # noqa: E402

N_IONS, SEED, STEPS, EVERY = 4000, 3, 10000, 100
KMAX = [6, 6, 6]


def run_config(coul):
    return salt_config(coulomb_type=coul, kmax=KMAX, temp_ctrl="NONE", energy_every=EVERY)


def energy_per_atom(t):
    n = N_IONS
    return 0.5 * t["ke_sum"] / n + t["e_lj"] + t["e_coul"] + t["e_long"]


def drift_stats(steps, e, dt):
    t = np.asarray(steps, float) * dt
    e = np.asarray(e, float)
    slope = float(np.polyfit(t, e, 1)[0])
    spread = float(np.abs(e - e[0]).max() / abs(e[0]))
    return slope, spread


def main():
    oracle.set_threads(int(os.environ.get("OMP_NUM_THREADS", "8")))
    s = make_salt(n_ions=N_IONS, seed=SEED)
    out = {"workload": f"salt {N_IONS} ions seed {SEED}, rc 4, alpha 0.64, dt 0.001, temp_ctrl NONE",
           "steps": STEPS, "every": EVERY, "kmax": KMAX, "runs": {}}
    for coul in ("EWALD", "RBE"):
        run = run_config(coul)
        o = oracle.OracleSystem(s, run)
        assert o.md_init() == 0
        steps, e = [0], [energy_per_atom(o.thermo())]
        t0 = time.time()
        for k in range(STEPS // EVERY):
            assert o.md_step(EVERY) == 0
            th = o.thermo()
            assert th["energy_step"] == th["step"]
            steps.append(th["step"])
            e.append(energy_per_atom(th))
        slope, spread = drift_stats(steps, e, run.dt)
        out["runs"][coul] = {"steps": steps, "energy": e, "slope": slope, "spread": spread,
                             "final_temperature": th["temperature"]}
        print(f"{coul}: slope {slope:.4e} per atom per tau, spread {spread:.3e}, {time.time() - t0:.0f} s", flush=True)
    with open(os.path.join(HERE, "nve_drift_salt4000.json"), "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
