"""Benchmark of the MI355X force path: MD steps of one synthetic system on one GPU.

Default workload (N=1): rigid 3-site water, 10,000 molecules (30,000 atoms),
1.0 g/cm^3 (L = 66.88 A), real-space cut 10 A, beta = 0.3 /A (alpha = 0.09),
RBE with P = 200 k-vectors per step, SHAKE, Nose-Hoover NVT 300 K, dt = 2 fs,
FP64 on the GPU.  A "step" is one full MD step (NHC half steps, kick, drift,
cell list, SHAKE, neighbour lists, real-space LJ + erfc, RBE structure factor
and force, bonded terms, kick, RATTLE) with all state resident in HBM.

Prints one JSON line; `value` is the water line's atom-steps/s.  Extra objects:
  roofline          the dominant kernel (most time per step) by algorithmic
                    work per launch over its mean duration (HIP events on the
                    library's streams): 43 flop per accepted pair for the pair
                    kernel (rejected candidates and skin entries not counted),
                    39 / 35 flop per (atom, k) term for the structure factor /
                    k-space force, bytes for the list build and binning
                    (kernel_work); peaks are MI355X's published FP64 vector
                    78.6 TFLOP/s (FP32 157.3 for mixed) and HBM 8 TB/s
  real_space_stage  list build + pair kernel per step against the 43 flop per
                    accepted pair convention, with the reference search's
                    candidate-test count beside it
  force_error       RBE k-space force against converged FP64 Ewald on the
                    run's final configuration, per step and averaged over the
                    first M k sets (M = 16, 64, 256, 1024), with the log-log
                    slope and the M that reaches 1e-3
  ensemble          512 salt/water systems x 500 steps: packs (mixed) against
                    the plain one-system-at-a-time FP64 loop (child process)
  served            Poisson job stream of the served workload at a scaled rate:
                    p50/p99 latency and atom-steps/s of the served loop and of
                    the plain loop on the same arrivals (child process)
  cpu_baseline      the CPU oracle (a C++ restatement of the reference
                    algorithms) on this host: OpenMP at up to 16 threads and
                    serial, on bounded samples of the same workload, with nproc
                    and the CPU model

roofline.traffic is the dominant kernel's HBM bytes per launch measured with
rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE (profiles/hbm_traffic.json, scaled by
the same counters' reading for a device copy of known size, see
--copy-calibrate); null when that kernel was not measured on this workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0
PAIR_FLOPS = 43.0  # per accepted pair (counted from each side): LJ + erfc-Coulomb force
TEST_FLOPS = 8.0  # per distance test of an atom against a stencil candidate


WORKLOAD_PRECISION = {"water": "fp64", "salt": "fp64", "water300k": "mixed", "ensemble": "mixed", "serve": "mixed"}
WORKLOAD_CPU_STEPS = {"water": 100, "salt": 100, "water300k": 10}
WORKLOAD_SKIN = {"water": 0.0, "salt": 0.0, "water300k": 0.0}
METRIC = "atom-steps/s + ns/day at force err \u22641e-3; served-ensemble p50/p99 vs plain loop"


def workload(name, precision, rbe_p=None):
    from rbmd_amd import make_salt, make_water, salt_config, water_config

    if name == "water":
        s = make_water(n_mol=10000, seed=1)
        r = water_config(precision=precision, **({"rbe_p": rbe_p} if rbe_p else {}))
        desc = ("rigid 3-site water, 10,000 molecules (30,000 atoms), 1.0 g/cm3 (L=66.88 A), rc 10 A, "
                "alpha 0.09 (beta 0.3/A), RBE P=200, SHAKE, NVT Nose-Hoover 300 K, dt 2 fs")
        dt_fs = 2.0
    elif name == "water300k":
        # P = 2000 keeps the per-step RBE error of the 30,000-atom box (it grows
        # like sqrt(N) at fixed P); at P = 200 the lattice start heats by
        # ~250 K per step and SHAKE fails within 10 steps, in the CPU oracle too
        P = rbe_p or 2000
        s = make_water(n_mol=100000, seed=1)
        r = water_config(precision=precision, rbe_p=P)
        desc = ("rigid 3-site water, 100,000 molecules (300,000 atoms), 1.0 g/cm3 (L=144.09 A), rc 10 A, "
                f"alpha 0.09 (beta 0.3/A), RBE P={P}, SHAKE, NVT Nose-Hoover 300 K, dt 2 fs")
        dt_fs = 2.0
    elif name == "salt":
        s = make_salt(n_ions=4000, seed=1)
        r = salt_config(precision=precision, **({"rbe_p": rbe_p} if rbe_p else {}))
        desc = ("charged-LJ salt, 4,000 ions, rho*=0.5 (L=20 sigma), rc 4 sigma, alpha 0.64 (beta 0.8), "
                "RBE P=100, NVT Nose-Hoover T*=1.5, dt 0.001 tau")
        dt_fs = None
    else:
        raise SystemExit(f"unknown workload {name}")
    return s, r, desc, dt_fs


def tile_choice(run, n_atoms):
    """the tile choices the system was created with (0: the library default)"""
    from rbmd_amd.abi import make_params

    p = make_params(run, n_atoms)
    return {"pair_lanes": p.pair_lanes, "kforce_lanes": p.kforce_lanes, "rho_kblock": p.rho_kblock}


def dtype_of(precision):
    return "f64" if precision == "fp64" else "f32 pair scalar and sin/cos, f64 vectors and sums"


def stencil_candidates(g):
    """sum over atoms of the atom counts of their 27 stencil cells (the distance
    tests of the reference's neighbour search)"""
    import numpy as np

    dims = g.geometry()[0]
    _, _, lower, upper = g.locator()
    cnt = (np.asarray(upper) - np.asarray(lower)).reshape(dims[2], dims[1], dims[0]).astype(np.int64)
    nb = np.zeros_like(cnt)
    for a in (-1, 0, 1):
        for b in (-1, 0, 1):
            for c in (-1, 0, 1):
                nb += np.roll(cnt, (c, b, a), axis=(0, 1, 2))
    return int((cnt * nb).sum())


def cpu_baseline(system, run, steps, threads):
    import oracle

    oracle.set_threads(threads)
    o = oracle.OracleSystem(system, run)
    o.md_init()
    o.md_step(1)
    t0 = time.perf_counter()
    o.md_step(steps)
    dt = time.perf_counter() - t0
    return system.n_atoms * steps / dt, dt


def ensemble_run(systems, steps, precision, mode="pack", streams=8, chunk=50, max_atoms=30000, workers=16):
    """Packed ensemble (one pack per kind, graph-replayed packed steps) against
    the plain one-system-at-a-time FP64 loop on the same systems.  Inputs are
    generated before anything is timed."""
    from rbmd_amd.serve import build_parallel, ensemble_specs, run_packed, run_plain

    specs = ensemble_specs(systems, steps, seed=2024, n_max=max_atoms)
    log(f"ensemble: {len(specs)} systems, {sum(s.atoms for s in specs)} atoms, {steps} steps each")
    t0 = time.perf_counter()
    inputs = build_parallel(specs, "fp64", workers=workers)
    log(f"ensemble inputs built in {time.perf_counter() - t0:.1f} s")
    ran = lambda i: log(f"plain loop: {i} systems done")
    plain = run_plain(specs, inputs=inputs, progress=ran)  # the plain loop is FP64 by definition
    log(f"plain loop {plain['atom_steps_per_s']:.4g} atom-steps/s")
    if precision != "fp64":
        inputs = [(s, type(r)(**{**r.__dict__, "precision": precision})) for s, r in inputs]
    packed = run_packed(specs, streams=streams, chunk=chunk, precision=precision, mode=mode, inputs=inputs,
                        progress=lambda m: log(f"{mode}: {m}"))
    log(f"{mode} {packed['atom_steps_per_s']:.4g} atom-steps/s")
    return {"workload": f"{len(specs)} independent systems, half charged-LJ salt, half rigid water, atoms "
                        f"log-uniform in [600, {max_atoms}] (seed 2024), {steps} steps each",
            "systems": len(specs), "atoms": sum(s.atoms for s in specs), "steps": steps,
            "precision": precision, "mode": mode, "chunk": chunk,
            "atom_steps_per_s": round(packed["atom_steps_per_s"], 1), "seconds": round(packed["seconds"], 3),
            "plain_loop_precision": "fp64",
            "plain_loop_atom_steps_per_s": round(plain["atom_steps_per_s"], 1),
            "plain_loop_seconds": round(plain["seconds"], 3),
            "speedup_vs_plain_loop": round(packed["atom_steps_per_s"] / plain["atom_steps_per_s"], 3)}


def served_run(rate, duration, precision, max_atoms=300000, max_steps=5000, mode="pack", streams=8, chunk=50,
               workers=16, plain=True):
    """Poisson job stream served with continuous batching (packs per kind,
    admission by footprint) and, on the same arrivals, by the plain loop (one
    FP64 system at a time in arrival order): p50 / p99 job latency and the
    aggregate atom-steps/s of each."""
    from rbmd_amd.serve import build_parallel, serve, serve_plain, stream_specs

    jobs = stream_specs(rate, duration, seed=77, n_max=max_atoms, steps_min=200, steps_max=max_steps)
    log(f"serve: {len(jobs)} jobs over {duration} s")
    inputs = build_parallel(jobs, precision, workers=workers)  # each job's input as it arrives
    say = lambda n: log(f"served {n} jobs")
    r = serve(jobs, streams=streams, chunk=chunk, prebuilt=dict(enumerate(inputs)), precision=precision, mode=mode,
              progress=say)
    offered = sum(sp.atoms * sp.steps for sp in jobs) / duration
    out = {"workload": f"Poisson arrivals {rate}/s for {duration} s (seed 77), salt or water 50/50, atoms "
                       f"log-uniform in [1000, {max_atoms}], steps uniform in [200, {max_steps}]",
           "precision": precision, "mode": mode, "chunk": chunk, "jobs": r["jobs"],
           "p50_latency_s": round(r["p50_latency_s"], 4), "p99_latency_s": round(r["p99_latency_s"], 4),
           "atom_steps_per_s": round(r["atom_steps_per_s"], 1), "offered_atom_steps_per_s": round(offered, 1),
           "wall_s": round(r["wall_s"], 3)}
    if plain:
        say = lambda n: log(f"plain loop served {n} jobs")
        q = serve_plain(jobs, prebuilt=dict(enumerate(inputs)), progress=say)
        out.update({"plain_loop_p50_latency_s": round(q["p50_latency_s"], 4),
                    "plain_loop_p99_latency_s": round(q["p99_latency_s"], 4),
                    "plain_loop_atom_steps_per_s": round(q["atom_steps_per_s"], 1),
                    "plain_loop_wall_s": round(q["wall_s"], 3)})
    return out


def bench_ensemble(args):
    e = ensemble_run(args.systems, args.ens_steps, args.precision, mode=args.ens_mode, streams=args.streams,
                     chunk=args.chunk, max_atoms=args.max_atoms)
    out = {"metric": "atom-steps/s (ensemble, aggregate)", "value": e["atom_steps_per_s"],
           "unit": "atom-steps/s", "n_gpus": 1, "steps": args.ens_steps, "warmup": 0,
           "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": dtype_of(args.precision), "data": "synthetic",
           "config": {"workload": e["workload"], "mode": args.ens_mode, "chunk": args.chunk,
                      "precision": args.precision, "plain_loop_precision": "fp64",
                      "hw_queues": int(os.environ.get("GPU_MAX_HW_QUEUES", "4"))},
           "plain_loop_atom_steps_per_s": e["plain_loop_atom_steps_per_s"],
           "speedup_vs_plain_loop": e["speedup_vs_plain_loop"], "leg": e}
    print(json.dumps(out))


def bench_serve(args):
    r = served_run(args.rate, args.duration, args.precision, max_atoms=args.max_atoms, max_steps=args.max_steps,
                   mode=args.ens_mode, streams=args.streams, chunk=args.chunk, plain=not args.no_plain_serve)
    out = {"metric": "served jobs: p50/p99 latency, aggregate atom-steps/s", "value": r["atom_steps_per_s"],
           "unit": "atom-steps/s", "n_gpus": 1, "steps": None, "warmup": 0, "ms_per_step": None,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": dtype_of(args.precision),
           "data": "synthetic", "config": {"workload": r["workload"], "mode": args.ens_mode, "chunk": args.chunk,
                                            "precision": args.precision},
           **{k: v for k, v in r.items() if k not in ("workload", "precision", "mode", "chunk")}, "leg": r}
    print(json.dumps(out))


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def kernel_work(name, system, run, entries, ncell):
    """Algorithmic work of one launch of a step kernel: (flops, bytes).
    pair: 43 flop per accepted pair (each side counted; rejected candidates
    and skin entries not counted); bytes: double4 (x, y, z, q) 32 + type 4 +
    order 4 + FP64 force 24 per atom, and 4 per list entry read.  List build:
    32 B position + 4 B count per atom, 4 B per entry written.  Binning (bin,
    scan, scatter, cell sort): 60 B/atom with FP64 positions + 8 B/cell.
    Structure factor 39 flop and k-space force 35 flop per (atom, k) term."""
    n = system.n_atoms
    P = run.rbe_p
    if name == "pair":
        return PAIR_FLOPS * entries, 64.0 * n + 4.0 * entries
    if name == "nbr_build":
        return None, 36.0 * n + 4.0 * entries
    if name in ("bin", "scan_scatter", "scan", "scatter", "cell_sort"):
        return None, 60.0 * n + 8.0 * ncell
    if name == "rho_partial":
        return 39.0 * n * P, 16.0 * n
    if name == "kforce":
        return 35.0 * n * P, 40.0 * n
    return None, None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--workload", default="water", choices=["water", "salt", "water300k", "ensemble", "serve"])
    ap.add_argument("--precision", choices=["fp64", "mixed"], default=None,
                    help="default: fp64 for water and salt, mixed for water300k, ensemble and serve")
    ap.add_argument("--skin", type=float, default=None,
                    help="Verlet-list skin of the MD steps (default: the workload's; 0: a list per step)")
    ap.add_argument("--systems", type=int, default=512, help="ensemble: number of systems (half salt, half water)")
    ap.add_argument("--ens-steps", type=int, default=500, help="ensemble: MD steps per system")
    ap.add_argument("--ens-mode", choices=["pack", "streams"], default="pack",
                    help="ensemble / serve: one pack per kind (one launch per kernel for all its systems) or "
                         "per-system step graphs spread over --streams contexts")
    ap.add_argument("--streams", type=int, default=8, help="ensemble / serve with --ens-mode streams: contexts")
    ap.add_argument("--chunk", type=int, default=50, help="ensemble / serve: graph steps per enqueue round")
    ap.add_argument("--rate", type=float, default=1.5, help="serve: Poisson arrivals per second")
    ap.add_argument("--duration", type=float, default=30.0, help="serve: seconds of arrivals")
    ap.add_argument("--max-atoms", type=int, default=None,
                    help="ensemble / serve: largest system (default 30,000 / 300,000)")
    ap.add_argument("--max-steps", type=int, default=5000, help="serve: most steps of a job")
    ap.add_argument("--no-ensemble", action="store_true", help="water / salt: skip the ensemble object")
    ap.add_argument("--no-serve", action="store_true", help="water / salt: skip the served-stream object")
    ap.add_argument("--no-plain-serve", action="store_true", help="serve: skip the plain-loop comparison")
    ap.add_argument("--leg-timeout", type=int, default=420, help="seconds allowed to the ensemble and served legs")
    ap.add_argument("--cpu-steps", type=int, default=None, help="OpenMP CPU baseline steps (default 100; 10 at 300k)")
    ap.add_argument("--serial-steps", type=int, default=None, help="serial CPU baseline steps (default 8; 2 at 300k)")
    ap.add_argument("--rbe-p", type=int, default=None, help="RBE k vectors per step (default: the workload's)")
    ap.add_argument("--hw-queues", type=int, default=None,
                    help="ensemble / serve: hardware queues HIP gives this process (GPU_MAX_HW_QUEUES, 1..32)")
    ap.add_argument("--error-sets", default="16,64,256,1024",
                    help="k-set counts M at which the averaged RBE-vs-Ewald force error is reported (0: skip)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-steps", type=int, default=50)
    ap.add_argument("--graph", action="store_true", help="replay a captured step graph (rbmd_md_run_graph)")
    ap.add_argument("--copy-calibrate", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device to device once (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.copy_calibrate:
        import torch

        n = args.copy_calibrate << 20
        a = torch.ones(n, dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n, "copy_bytes_written": n}))
        return
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; --gpus must be 1 (multi-GPU sharding is design-only)",
              file=sys.stderr)
        raise SystemExit(2)

    if args.hw_queues is not None:
        if not 1 <= args.hw_queues <= 32:
            raise SystemExit("--hw-queues must be in 1..32")
        os.environ["GPU_MAX_HW_QUEUES"] = str(args.hw_queues)  # read when HIP initialises, below
    from rbmd_amd.lib import Context, GpuSystem

    if args.precision is None:
        args.precision = WORKLOAD_PRECISION[args.workload]
    if args.cpu_steps is None:
        args.cpu_steps = WORKLOAD_CPU_STEPS.get(args.workload, 100)
    if args.serial_steps is None:
        args.serial_steps = 2 if args.workload == "water300k" else 8
    if args.workload == "ensemble":
        args.max_atoms = args.max_atoms or 30000
        return bench_ensemble(args)
    if args.workload == "serve":
        args.max_atoms = args.max_atoms or 300000
        return bench_serve(args)
    system, run, desc, dt_fs = workload(args.workload, args.precision, args.rbe_p)
    skin = WORKLOAD_SKIN.get(args.workload, 0.0) if args.skin is None else args.skin
    run.skin = skin
    if args.rbe_p and args.workload != "water300k":
        desc = desc.replace("RBE P=200", f"RBE P={args.rbe_p}").replace("RBE P=100", f"RBE P={args.rbe_p}")
    log(f"workload {args.workload}: {system.n_atoms} atoms, skin {skin}")
    ctx = Context(0)
    g = GpuSystem(ctx, system, run)
    g.md_init()
    run_steps = g.md_run_graph if args.graph else g.md_step
    run_steps(args.warmup)
    ctx.synchronize()
    t0 = time.perf_counter()
    run_steps(args.steps)
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed * 1e3 / args.steps
    value = system.n_atoms * args.steps / elapsed
    th_timed = g.thermo()
    log(f"timed {args.steps} steps in {elapsed:.3f} s, T {th_timed['temperature']:.2f}")

    # per-kernel mean durations with HIP events on the library's streams
    ctx.profile(True)
    ctx.profile_reset()
    g.md_step(args.profile_steps)
    rec = ctx.profile_records()
    ctx.profile(False)
    th = g.thermo()
    # the reference's per-step list on the final configuration: its entries are
    # the accepted pairs (each side counted) of the algorithmic pair work
    cands = stencil_candidates(g)
    g.build_neighbors(0)
    entries = g.neighbor_entries()
    dims = g.geometry()[0]
    ncell = int(dims[0] * dims[1] * dims[2])
    steps_prof = args.profile_steps
    kernels = {k: {"launches": n, "mean_us": 1e3 * ms / max(n, 1), "us_per_step": 1e3 * ms / steps_prof}
               for k, (n, ms) in rec.items()}
    dom = max(rec.items(), key=lambda kv: kv[1][1])[0]
    P = run.rbe_p

    def roof_of(name):
        k = kernels[name]
        flops, nbytes = kernel_work(name, system, run, entries, ncell)
        t = k["mean_us"] * 1e-6
        if flops is not None:
            if args.precision == "mixed" and name in ("pair", "rho_partial", "kforce"):
                peak, note = FP32_PEAK_TFLOPS, "MI355X FP32 vector peak (FP32 scalar math, FP64 vectors and sums)"
            else:
                peak, note = FP64_PEAK_TFLOPS, "MI355X FP64 vector peak"
            a = flops / t / 1e12
            r = {"kernel": name, "bound": "valu", "achieved": round(a, 3), "peak": peak, "unit": "TFLOP/s",
                 "frac": round(a / peak, 4), "work_per_launch": flops, "peak_note": note}
            if nbytes is not None:
                r["hbm_gbs_algorithmic"] = round(nbytes / t / 1e9, 1)
            return r
        a = nbytes / t / 1e9
        return {"kernel": name, "bound": "hbm", "achieved": round(a, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(a / HBM_PEAK_GBS, 4), "bytes_per_launch": nbytes}

    traffic = None
    tpath = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tj = json.load(f)
        if tj.get("bench_args") == traffic_key(args.workload, args.precision, skin):  # measured on this config only
            tt = tj.get("kernels", {}).get(dom)
            if tt:
                traffic = tt["bytes_per_launch"]
    roof = None
    if kernel_work(dom, system, run, entries, ncell) != (None, None):
        roof = roof_of(dom)
        roof["traffic"] = traffic
    kroofs = {k: roof_of(k) for k in kernels if kernel_work(k, system, run, entries, ncell) != (None, None)}
    # the real-space stage (list build + pair kernel per step) by the 43 flop per
    # accepted pair convention; the candidate-test count of the reference's
    # 27-cell search is kept beside it
    rs_us = sum(kernels[k]["us_per_step"] for k in ("pair", "nbr_build") if k in kernels)
    real_space = None
    if rs_us > 0:
        a = PAIR_FLOPS * entries / (rs_us * 1e-6) / 1e12
        real_space = {"us_per_step": round(rs_us, 2), "achieved": round(a, 3), "unit": "TFLOP/s",
                      "peak": FP64_PEAK_TFLOPS if args.precision == "fp64" else FP32_PEAK_TFLOPS,
                      "frac": round(a / (FP64_PEAK_TFLOPS if args.precision == "fp64" else FP32_PEAK_TFLOPS), 4),
                      "accepted_pairs": entries, "stencil_candidates": cands,
                      "candidate_test_flops_per_step": TEST_FLOPS * cands}
    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype_of(args.precision),
        "data": "synthetic",
        "config": {"workload": desc, "atoms": system.n_atoms, "precision": args.precision, "rbe_p": P,
                   "cut_off": run.cut_off, "skin": skin,
                   "list": (f"Verlet list to rc + {skin} kept across steps, rebuilt on the device when an atom "
                            f"moves more than skin/2; pairs tested against rc every step") if skin > 0 else
                           "Verlet list rebuilt every step (the reference's schedule)",
                   "energy_every": run.energy_every,
                   "energy_note": "kmax-grid Ewald energy not formed per step (the reference forms it every "
                                  "step for its thermo output; the trajectory does not depend on it)",
                   "thermo_sums": "virial and energy sums on the last step of each call only; kinetic energy "
                                  "every step (the thermostat needs it)",
                   "graph": bool(args.graph), "tiles": tile_choice(run, system.n_atoms)},
        "ns_per_day": round((args.steps / elapsed) * dt_fs * 86400 / 1e6, 3) if dt_fs else None,
        "roofline": roof,
        "real_space_stage": real_space,
        "kernels_roofline": kroofs,
        "kernels_us": {k: round(v["mean_us"], 2) for k, v in kernels.items()},
        "kernels_us_per_step": {k: round(v["us_per_step"], 2) for k, v in kernels.items()},
        "list_builds": th.get("list_builds"),
        "misfit_atoms": th["misfit_atoms"],
        "temperature_target": run.temperature[1],
        "final_temperature": th_timed["temperature"],
        "steps_run": args.warmup + args.steps,
        "status_flags": th["status_flags"],
    }
    sets = [int(m) for m in str(args.error_sets).split(",") if m.strip()]
    if sets and max(sets) > 0:
        from rbmd_amd.accuracy import precision_error, rbe_method_error

        # errors on the configuration the run ended on
        x, v, _, _ = g.state()
        cur = type(system)(**{**system.__dict__, "x": x, "v": v})
        fe = {}
        if run.coulomb_type == "RBE":
            t1 = time.perf_counter()
            me = rbe_method_error(ctx, cur, run, n_sets=max(sets), checkpoints=sets)
            log(f"force error over {max(sets)} k sets in {time.perf_counter() - t1:.1f} s")
            fe.update({"rbe_vs_ewald_per_step": round(me["per_step_rel_total"], 5),
                       "rbe_vs_ewald_averaged_by_m": {m: round(e, 6) for m, e in
                                                      me.get("averaged_rel_total_by_m", {}).items()},
                       "log_slope": round(me["log_slope"], 4) if "log_slope" in me else None,
                       "m_for_1e-3": round(me["m_for_1e-3"]) if me.get("m_for_1e-3") else None,
                       "ewald_kmax": me["kmax"]})
        if args.precision == "mixed":
            pe = precision_error(ctx, cur, run)
            fe.update({"mixed_vs_fp64_total": pe.get("total_rel_rms"), "mixed_vs_fp64_real": pe["real_rel_rms"],
                       "mixed_vs_fp64_rbe": pe.get("rbe_rel_rms"), "mixed_vs_fp64_e_lj": pe["e_lj_rel"],
                       "mixed_vs_fp64_e_coul": pe["e_coul_rel"]})
        fe["definition"] = ("relative RMS over atoms against the FP64 reference's nonbonded force "
                            "(real-space LJ+erfc + k space); RBE vs converged FP64 Ewald at the same alpha, "
                            "per step (one k set) and for the average over the first M k sets; m_for_1e-3 "
                            "extrapolates the log-log fit to an averaged error of 1e-3")
        out["force_error"] = fe
    g.close()
    ctx.close()
    extra_ok = args.workload in ("water", "salt")
    if extra_ok and not args.no_ensemble:
        out["ensemble"] = child_leg(["--workload", "ensemble", "--systems", str(args.systems), "--ens-steps",
                                     str(args.ens_steps), "--chunk", str(args.chunk)], args.leg_timeout)
    if extra_ok and not args.no_serve:
        out["served"] = child_leg(["--workload", "serve", "--rate", str(args.rate), "--duration",
                                   str(args.duration), "--max-steps", str(args.max_steps), "--chunk",
                                   str(args.chunk)] + (["--no-plain-serve"] if args.no_plain_serve else []),
                                  args.leg_timeout)
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None
        log(f"CPU baseline: {args.cpu_steps} steps on {threads} threads")
        cpu_value, cpu_time = cpu_baseline(system, run, args.cpu_steps, threads)
        log(f"CPU baseline: {args.serial_steps} steps on 1 thread")
        ser_value, ser_time = cpu_baseline(system, run, args.serial_steps, 1)
        out["cpu_baseline"] = {"value": round(cpu_value, 1), "unit": "atom-steps/s", "cores": threads,
                               "kind": "port",
                               "sample": f"{args.cpu_steps} MD steps of the same workload after init + 1 step "
                                         f"({cpu_time:.1f} s, OpenMP FP64 oracle, per-step list)",
                               "serial": {"value": round(ser_value, 1), "cores": 1,
                                          "sample": f"{args.serial_steps} MD steps after init + 1 step "
                                                    f"({ser_time:.1f} s)"},
                               "nproc": os.cpu_count(), "affinity_cpus": aff, "cpu_model": cpu_model()}
    print(json.dumps(out))


def child_leg(argv, timeout):
    """one workload of the line (ensemble, served stream) in a child process of
    its own (fresh device state; its progress lines go to this stderr); its
    measured object, or the reason it has none"""
    import subprocess

    cmd = [sys.executable, "-u", os.path.abspath(__file__)] + argv
    log("leg: " + " ".join(argv))
    try:
        r = subprocess.run(cmd, stdout=subprocess.PIPE, text=True, timeout=timeout)
    except subprocess.TimeoutExpired:
        return {"error": f"no result within {timeout} s", "command": " ".join(argv)}
    lines = [ln for ln in r.stdout.splitlines() if ln.startswith("{")]
    if r.returncode != 0 or not lines:
        return {"error": f"exit status {r.returncode}", "command": " ".join(argv)}
    leg = json.loads(lines[-1]).get("leg", {})
    leg["command"] = "python bench.py " + " ".join(argv)
    return leg


def kernel_label(full):
    """the bench's label (ctx.profile_records) of a kernel's full name in a rocprofv3 database"""
    name = full.replace("void ", "").replace("rbmd::", "").split("(")[0]
    base, _, targs = name.partition("<")
    targs = [t.strip() for t in targs.rstrip(">").split(",")] if targs else []
    if base == "k_pair_verlet" and len(targs) > 1:  # <MODE, THERMO, ...>
        return "pair_thermo" if targs[1] == "true" else "pair"
    if base == "k_kforce" and targs:  # <THERMO, ...>
        return "kforce_thermo" if targs[0] == "true" else "kforce"
    if base == "k_nbr_build_verlet":
        return "nbr_build"
    return base[2:] if base.startswith("k_") else base


def pmc_means(run_dir, counter):
    """per-kernel mean over dispatches of one counter (summed over its instances) in a rocprofv3 --pmc run"""
    import collections
    import glob
    import sqlite3

    dbs = glob.glob(os.path.join(run_dir, "**", "*.db"), recursive=True)
    if len(dbs) != 1:
        raise SystemExit(f"{run_dir}: expected one rocprofv3 database, found {len(dbs)}")
    con = sqlite3.connect(dbs[0])
    per = collections.defaultdict(float)
    for kname, disp, v in con.execute("select kernel_name, dispatch_id, value from counters_collection "
                                      "where counter_name = ?", (counter,)):
        per[(kname, disp)] += v
    con.close()
    acc = collections.defaultdict(list)
    for (kname, _), v in per.items():
        acc[kname].append(v)
    return {k: (sum(v) / len(v), len(v)) for k, v in acc.items()}


def traffic_summary(argv):
    """write profiles/hbm_traffic.json from four rocprofv3 --pmc runs: FETCH_SIZE and WRITE_SIZE of a
    bench run, then of --copy-calibrate MIB (argv: fetch_dir write_dir copy_fetch_dir copy_write_dir MIB
    bench_args...)"""
    hf, hw, cf, cw, mib = argv[:5]
    bench_args = " ".join(argv[5:])
    copy_kib = int(mib) * 1024.0
    cfm, cwm = pmc_means(cf, "FETCH_SIZE"), pmc_means(cw, "WRITE_SIZE")
    ck = max(cfm, key=lambda k: cfm[k][0])  # the copy kernel: the largest reader of the calibration run
    scale_f, scale_w = copy_kib / cfm[ck][0], copy_kib / cwm[ck][0]
    fm, wm = pmc_means(hf, "FETCH_SIZE"), pmc_means(hw, "WRITE_SIZE")
    kernels = {}
    for full, (f, n) in sorted(fm.items(), key=lambda kv: -kv[1][0]):
        if full not in wm or "rbmd::" not in full:
            continue
        w = wm[full][0]
        kernels[kernel_label(full)] = {"kernel": full, "dispatches": n, "fetch_kib": round(f, 1),
                                       "write_kib": round(w, 1),
                                       "bytes_per_launch": round((f * scale_f + w * scale_w) * 1024)}
    out = {"bench_args": bench_args,
           "commands": [f"rocprofv3 --pmc FETCH_SIZE -d {hf} -- python3 bench.py ... {bench_args}",
                        f"rocprofv3 --pmc WRITE_SIZE -d {hw} -- python3 bench.py ... {bench_args}",
                        f"rocprofv3 --pmc FETCH_SIZE -d {cf} -- python3 bench.py --copy-calibrate {mib}",
                        f"rocprofv3 --pmc WRITE_SIZE -d {cw} -- python3 bench.py --copy-calibrate {mib}",
                        "python3 bench.py --traffic-summary " + " ".join(argv)],
           "note": ("per-dispatch means; bytes_per_launch = fetch_kib*scale_fetch*1024 + write_kib*scale_write*1024;"
                    f" the scales make the counters read {mib} MiB for a {mib} MiB device copy ({ck}); the _thermo"
                    " entries are the last step of each call (virial and energy sums)"),
           "scale_fetch": round(scale_f, 4), "scale_write": round(scale_w, 4), "kernels": kernels}
    with open(os.path.join(ROOT, "profiles", "hbm_traffic.json"), "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")
    print(json.dumps({k: v["bytes_per_launch"] for k, v in kernels.items()}))


def traffic_key(workload, precision, skin):
    """the bench arguments an HBM-traffic record was measured with"""
    return f"--workload {workload} --precision {precision} --skin {skin:g}"


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "--traffic-summary":
        traffic_summary(sys.argv[2:])
    else:
        main()
