"""Tile autotuner: measures the kernels' tile choices on the GPU and writes
rbmd_amd/tiles.json, which the host layer applies (rbmd_amd.tiles).

    python -m rbmd_amd.tune [--out rbmd_amd/tiles.json] [--sizes 1000,3000,...] [--steps 20] [--step-steps 200]

For each kind (charged-LJ salt, rigid water), precision (fp64, mixed) and
system size it builds the workload of that shape (water above 30,000 atoms
with P = 200 N / 30,000 k vectors, as the served loop runs it) and times, with
HIP events on the library's stream, every option of

  pair_lanes    lanes per atom of the pair kernel          4, 8, 16
  kforce_lanes  wavefronts splitting the k vectors of the k-space force kernel  2, 4, 8
  rho_kblock    k vectors per lane of the structure-factor kernel               1, 2, 4

(option i of all three in run i: the kernels are timed separately, in the
library's kernel-profiling mode, on one stream).  The pair lanes are chosen
per size (a per-system runtime value) by the pair kernel's time.

The two k-space choices are made per kind and precision (the kernels of a
pack take them as template parameters, so all systems of a pack share them)
by the wall time of whole steps (--step-steps steps, 200 by default: the
options differ by 1-3%, as much as 20-step timings scatter), one run per
option of each with the other tiles at the library's defaults: in a step the k-space kernels run on a
second stream beside the list build and pair kernel, and there their
isolated times mislead (water, 300,000 atoms, mixed: two k vectors per lane
times 322 us against 431 us for one alone, but the step is 3% slower).  The
rule is the smallest sum over sizes of each option's time relative to the
best option at that size (every size counts alike).  The table keeps the raw
timings ("step_<tile>": step times in us), and
`python -m rbmd_amd.tune --rechoose tiles.json` recomputes the choices from
them.
"""
import argparse
import json
import math
import os
import sys
import time

OPTIONS = {"pair_lanes": (4, 8, 16), "kforce_lanes": (2, 4, 8), "rho_kblock": (1, 2, 4)}
KERNEL = {"pair_lanes": "pair", "kforce_lanes": "kforce", "rho_kblock": "rho_partial"}
KINDS = {"salt": "LJ/CUT/COUL/LONG", "water": "CVFF"}
K_TILES = ("kforce_lanes", "rho_kblock")  # chosen by step time


def measure(ctx, kind, precision, atoms, steps, step_steps, log):
    from .lib import GpuSystem
    from .serve import JobSpec, build

    s, r = build(JobSpec(kind, atoms, steps, seed=7), precision)
    out = {key: [] for key in OPTIONS}
    for i in range(3):
        run = type(r)(**{**r.__dict__, **{key: OPTIONS[key][i] for key in OPTIONS}})
        g = GpuSystem(ctx, s, run)
        g.md_init()
        g.md_step(3)
        ctx.synchronize()
        ctx.profile(True)
        ctx.profile_reset()
        g.md_step(steps)
        rec = ctx.profile_records()
        ctx.profile(False)
        g.close()
        for key in OPTIONS:
            n, ms = rec.get(KERNEL[key], (0, 0.0))
            out[key].append(1e3 * ms / max(n, 1))
        log(f"[tune] {kind} {precision} n={s.n_atoms} option {i}: " +
            ", ".join(f"{KERNEL[k]} {out[k][-1]:.1f} us" for k in OPTIONS))
    for key in K_TILES:
        out["step_" + key] = []
        for i in range(3):
            run = type(r)(**{**r.__dict__, "tiles": "default", "pair_lanes": 0, "kforce_lanes": 0, "rho_kblock": 0,
                             key: OPTIONS[key][i]})
            g = GpuSystem(ctx, s, run)
            g.md_init()
            g.md_step(3)
            ctx.synchronize()
            t0 = time.perf_counter()
            g.md_step(step_steps)
            ctx.synchronize()
            out["step_" + key].append(1e6 * (time.perf_counter() - t0) / step_steps)
            g.close()
        log(f"[tune] {kind} {precision} n={s.n_atoms} step us by {key}: " +
            ", ".join(f"{v:.1f}" for v in out["step_" + key]))
    return s.n_atoms, out


def choose(per_size):
    """tile choices from [(atoms, {key: [time of option 0, 1, 2]})] sorted by atoms"""
    entry = {}
    for key in K_TILES:
        times = [t.get("step_" + key, t[key]) for _, t in per_size]  # step times where measured
        score = [sum(v[i] / min(v) for v in times) for i in range(3)]
        entry[key] = OPTIONS[key][min(range(3), key=lambda i: score[i])]
    lanes = []
    for k, (n, t) in enumerate(per_size):
        best = OPTIONS["pair_lanes"][min(range(3), key=lambda i: t["pair_lanes"][i])]
        # bucket edge: geometric midpoint to the next measured size
        edge = int(math.sqrt(n * per_size[k + 1][0])) if k + 1 < len(per_size) else 2 ** 31 - 1
        lanes.append([edge, best])
    entry["pair_lanes"] = lanes
    entry["measured_us"] = {str(n): t for n, t in per_size}
    return entry


def tune(sizes, steps, step_steps, log):
    from .lib import Context

    ctx = Context(0)
    table = {"generated_by": "python -m rbmd_amd.tune", "sizes": sizes, "steps": steps, "step_steps": step_steps,
             "date": time.strftime("%Y-%m-%d"), "kinds": {}}
    for kind, ff in KINDS.items():
        for precision in ("fp64", "mixed"):
            entry = choose([measure(ctx, kind, precision, n, steps, step_steps, log) for n in sizes])
            table["kinds"].setdefault(ff, {})[precision] = entry
            log(f"[tune] {kind} {precision}: {json.dumps({k: v for k, v in entry.items() if k != 'measured_us'})}")
    ctx.close()
    return table


def rechoose(table):
    """recompute every choice of a table from its measured timings"""
    for ff, by_prec in table["kinds"].items():
        for precision, entry in by_prec.items():
            per_size = sorted(((int(n), t) for n, t in entry["measured_us"].items()), key=lambda v: v[0])
            by_prec[precision] = choose(per_size)
    return table


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(os.path.dirname(os.path.abspath(__file__)), "tiles.json"))
    ap.add_argument("--sizes", default="1000,3000,10000,30000,100000,300000")
    ap.add_argument("--steps", type=int, default=20, help="profiled steps per kernel timing")
    ap.add_argument("--step-steps", type=int, default=200, help="steps per step-time run (k-space tiles)")
    ap.add_argument("--rechoose", metavar="TABLE", help="recompute the choices of an existing table (no GPU)")
    args = ap.parse_args(argv)
    if args.rechoose:
        with open(args.rechoose) as f:
            table = rechoose(json.load(f))
    else:
        sizes = [int(v) for v in args.sizes.split(",")]
        table = tune(sizes, args.steps, args.step_steps, lambda m: print(m, file=sys.stderr, flush=True))
    with open(args.out, "w") as f:
        json.dump(table, f, indent=1)
    print(args.out)


if __name__ == "__main__":
    main()
