"""`rbmd -j in.json` on the MI355X path:

    python -m rbmd_amd.run -j in.json [--out DIR] [--steps N] [--precision fp64|mixed]

Reads the reference's JSON (BOM, comments, trailing commas) and its
init_configuration (read_data file next to the JSON, or the in-build lattice),
runs the executioner loop of Executioner::Execute
(framework/executioner/Executioner.cpp:62-129) on the GPU and writes the
reference's output files into DIR:

  energy.rbmd       ThermoOutput (framework/output/ThermoOutput.cpp:223-291)
  temperature.rbmd  TempOutput (framework/output/TempOutput.cpp:60-90)
  rbmd.trj          TrajectoryOutput (framework/output/TrajectoryOutput.cpp:44-80)

The loop runs num_steps + 1 steps; before step k (k = 0 .. num_steps) the
outputs see the state after k steps (k = 0: the state after Init).  thermo_out
rows are written at k % interval == 0 and at k = num_steps, trajectory frames
at k >= 1 with k % interval == 0.  Numbers are printed as C++ streams print
them by default (6 significant digits).  Time columns are wall-clock seconds,
as in the reference.  The other outputs (rdf_out, msd_out, vacf_out) are not
part of this build; they are reported and skipped.
"""
import argparse
import os
import sys
import time

from .config import load_json, run_config_from_json
from .system import make_inbuild, read_data


def cfmt(v) -> str:
    """a number as `std::ostream << v` prints it with the default precision"""
    if isinstance(v, (int,)) and not isinstance(v, bool):
        return str(v)
    return "%g" % float(v)


THERMO_HEADER = ("Step Time VanderWaalsEnergy NearCoulombicEnergy FarCoulombicEnergy Residual KinticEnergy "
                 "PotentialEnergy TotalEnergy CumulativeTime")
THERMO_HEADER_MOL = " NonBonEnergy BondEnergy AngleEnergy DihedralEnergy SpecialEnergy"


def energies(force_field: str, n_atoms: int, t: dict) -> dict:
    """ThermoOutput::ComputePotentialEnergy (ThermoOutput.cpp:140-180) from rbmd_md_thermo values"""
    e = {"lj": t["e_lj"], "near": 0.0, "far": 0.0, "nonbond": 0.0, "bond": t["e_bond"], "angle": t["e_angle"]}
    if force_field == "LJ/CUT":
        e["pe"] = t["e_lj"]
    elif force_field == "LJ/CUT/COUL/LONG":
        e["near"], e["far"] = t["e_coul"], t["e_long"]
        e["pe"] = e["lj"] + e["far"] + e["near"]
    else:  # CVFF
        e["near"], e["far"] = t["e_coul"], t["e_long"]
        e["nonbond"] = e["lj"] + e["far"] + e["near"]
        e["pe"] = e["nonbond"] + e["bond"] + e["angle"]
    e["ke"] = t["ke_sum"] / n_atoms * 0.5
    e["total"] = e["pe"] + e["ke"]
    return e


def thermo_row(step, wall, e, residual, cumulative, molecular) -> str:
    cols = [step, wall, e["lj"], e["near"], e["far"], residual, e["ke"], e["pe"], e["total"], cumulative]
    row = " ".join(cfmt(c) for c in cols) + " "
    if molecular:
        row += " ".join(cfmt(c) for c in (e["nonbond"], e["bond"], e["angle"], 0.0, 0.0))
    return row


def trajectory_frame(step, lo, hi, types, x) -> str:
    out = [f"ITEM: TIMESTEP\n{step}\nITEM: NUMBER OF ATOMS\n{len(x)}\nITEM: BOX BOUNDS pp pp pp\n"]
    for d in range(3):
        out.append(f"{cfmt(lo[d])} {cfmt(hi[d])}\n")
    out.append("ITEM: ATOMS id type x y z\n")
    out.extend(f"{i + 1} {int(types[i]) + 1} {cfmt(p[0])} {cfmt(p[1])} {cfmt(p[2])}\n" for i, p in enumerate(x))
    return "".join(out)


def load_system(doc: dict, json_dir: str, run, seed: int = 1):
    init = doc.get("init_configuration", {}) or {}
    if "read_data" in init:
        rd = init["read_data"]
        path = rd["file"] if os.path.isabs(rd["file"]) else os.path.join(json_dir, rd["file"])
        return read_data(path, rd.get("atom_style", "full"), rd["unit"], special_bonds=run.special_bonds)
    ib = init["inbuild"]
    return make_inbuild(ib["dims"], ib["x_range"], ib["y_range"], ib["z_range"], ib.get("velocity_type", "GAUSS"), seed)


def run_md(sim, system, run, doc: dict, out_dir: str, num_steps: int, log=None) -> dict:
    """The executioner loop over `sim` (a GpuSystem, or any object with the
    same md_init / md_step / md_energy / thermo / state methods) with the
    reference's outputs.  Returns the last thermo row's values."""
    outs = doc.get("outputs", {}) or {}
    thermo_iv = int(outs["thermo_out"].get("interval", 1)) if "thermo_out" in outs else 0
    traj_iv = int(outs["trajectory_out"].get("interval", 1)) if "trajectory_out" in outs else 0
    for other in sorted(set(outs) - {"thermo_out", "trajectory_out"}):
        if log:
            log(f"[rbmd] {other}: not part of this build, skipped")
    os.makedirs(out_dir, exist_ok=True)
    fe = open(os.path.join(out_dir, "energy.rbmd"), "w") if thermo_iv else None
    ft = open(os.path.join(out_dir, "temperature.rbmd"), "w") if thermo_iv else None
    fj = open(os.path.join(out_dir, "rbmd.trj"), "w") if traj_iv else None
    molecular = not run.inbuild
    has_long = run.force_field != "LJ/CUT" and run.coulomb_type != "NONE"
    n = system.n_atoms
    lo, hi = system.lo, system.hi

    def thermo_due(k):
        return thermo_iv and k <= num_steps and (k == num_steps or k % thermo_iv == 0)

    def traj_due(k):
        return traj_iv and 1 <= k <= num_steps and k % traj_iv == 0

    residual_on = molecular and run.force_field == "CVFF"  # _potential_energy_avr is set for CVFF only

    def nonbond_now():
        if has_long:
            sim.md_energy()
        return energies(run.force_field, n, sim.thermo())["nonbond"]

    sim.md_init()
    if fe:
        fe.write(THERMO_HEADER + (THERMO_HEADER_MOL if molecular else "") + "\n")
        ft.write("Step Temperature Pressure\n")
    t_start = time.perf_counter()
    t_last = t_start
    prev_nonbond = None  # non-bonded energy of the step before the next thermo row (Residual)
    last = {}
    k = 0
    while k <= num_steps:
        if thermo_due(k):
            if has_long:
                sim.md_energy()  # ComputeEwaldEnergy of this step
            t = sim.thermo()
            e = energies(run.force_field, n, t)
            now = time.perf_counter()
            residual = e["nonbond"] - prev_nonbond if (residual_on and k > 0 and prev_nonbond is not None) else 0.0
            fe.write(thermo_row(k, now - t_last, e, residual, now - t_start, molecular) + "\n")
            ft.write(f"{k} {cfmt(t['temperature'])} {cfmt(t['pressure'])}\n")
            t_last = now
            last = {"step": k, "temperature": t["temperature"], "pressure": t["pressure"], **e}
            if log:
                log(f"[rbmd] step {k} T {cfmt(t['temperature'])} P {cfmt(t['pressure'])} "
                    f"PE {cfmt(e['pe'])} TE {cfmt(e['total'])}")
        if traj_due(k):
            fj.write(trajectory_frame(k, lo, hi, system.type, sim.state()[0]))
        # advance to the next step with an output (or past the last step)
        nxt = k + 1
        while nxt <= num_steps and not (thermo_due(nxt) or traj_due(nxt)):
            nxt += 1
        if residual_on and thermo_due(nxt):
            if nxt - 1 > k:
                sim.md_step(nxt - 1 - k)
                prev_nonbond = nonbond_now()
            elif thermo_due(k):
                prev_nonbond = last["nonbond"]
            else:
                prev_nonbond = nonbond_now()
            sim.md_step(1)
        else:
            sim.md_step(nxt - k)
        k = nxt
    for f in (fe, ft, fj):
        if f:
            f.close()
    return last


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("-j", "--json", required=True)
    ap.add_argument("--out", default=".", help="directory for energy.rbmd, temperature.rbmd, rbmd.trj")
    ap.add_argument("--steps", type=int, default=None, help="override execution.num_steps")
    ap.add_argument("--precision", choices=["fp64", "mixed"], default=None)
    ap.add_argument("--seed", type=int, default=1, help="k-vector sampler seed and in-build velocities")
    args = ap.parse_args(argv)
    doc = load_json(args.json)
    run = run_config_from_json(doc)
    if args.precision:
        run.precision = args.precision
    run.seed = args.seed
    run.energy_every = 0  # the loop evaluates the Ewald energy where an output needs it
    system = load_system(doc, os.path.dirname(os.path.abspath(args.json)), run, args.seed)
    from .lib import Context, GpuSystem

    ctx = Context(0)
    g = GpuSystem(ctx, system, run)
    steps = run.num_steps if args.steps is None else args.steps
    say = lambda m: print(m, file=sys.stderr, flush=True)
    say(f"[rbmd] {system.n_atoms} atoms, {run.force_field}, {run.neighbor_type}, {run.coulomb_type}, "
        f"{steps} + 1 steps, {run.precision}")
    run_md(g, system, run, doc, args.out, steps, log=say)
    g.close()
    ctx.close()


if __name__ == "__main__":
    main()
