"""Benchmark of the MI355X force path: MD steps of one synthetic system on one GPU.

Default workload (N=1): rigid 3-site water, 10,000 molecules (30,000 atoms),
1.0 g/cm^3 (L = 66.88 A), real-space cut 10 A, beta = 0.3 /A (alpha = 0.09),
RBE with P = 200 k-vectors per step, SHAKE, Nose-Hoover NVT 300 K, dt = 2 fs,
FP64 on the GPU.  A "step" is one full MD step (NHC half steps, kick, drift,
cell list, SHAKE, neighbour lists, real-space LJ + erfc, RBE structure factor
and force, bonded terms, kick, RATTLE) with all state resident in HBM.

Prints one JSON line.  Extra objects:
  roofline      dominant kernel: algorithmic FP64 flops per launch / its mean
                duration measured with HIP events on the library's stream,
                against the FP64 vector peak (78.6 TFLOP/s).  Algorithmic work
                is the reference algorithm's: 8 flops per distance test of an
                atom against every atom of its 27 stencil cells, 43 per
                accepted pair (pair term, counted from each side)
  cpu_baseline  the CPU oracle (a C++ restatement of the reference
                algorithms) timed on this host's cores on a bounded sample of
                the same workload (--cpu-steps MD steps, about 10 s)

roofline.traffic is the dominant kernel's HBM bytes per launch measured with
rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE (profiles/hbm_traffic.json, scaled by
the same counters' reading for a device copy of known size, see
--copy-calibrate); null when that kernel was not measured.
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


def bench_ensemble(args):
    """Packed ensemble (graph steps on several streams) against the plain
    one-system-at-a-time FP64 loop on the same systems."""
    from rbmd_amd.serve import build_all, ensemble_specs, run_packed, run_plain

    specs = ensemble_specs(args.systems, args.ens_steps, seed=2024, n_max=args.max_atoms)
    print(f"[bench] ensemble: {len(specs)} systems, {sum(s.atoms for s in specs)} atoms, "
          f"{args.ens_steps} steps each", file=sys.stderr, flush=True)
    say = lambda i: print(f"[bench] built {i} systems", file=sys.stderr, flush=True)
    inputs = build_all(specs, "fp64", progress=say)
    ran = lambda i: print(f"[bench] plain loop: {i} systems done", file=sys.stderr, flush=True)
    plain = run_plain(specs, inputs=inputs, progress=ran)  # the plain loop is FP64 by definition
    print(f"[bench] plain loop {plain['atom_steps_per_s']:.4g} atom-steps/s", file=sys.stderr, flush=True)
    if args.precision != "fp64":
        inputs = [(s, type(r)(**{**r.__dict__, "precision": args.precision})) for s, r in inputs]
    packed = run_packed(specs, streams=args.streams, chunk=args.chunk, precision=args.precision, mode=args.ens_mode,
                        inputs=inputs)
    print(f"[bench] {args.ens_mode} {packed['atom_steps_per_s']:.4g} atom-steps/s", file=sys.stderr, flush=True)
    out = {"metric": "atom-steps/s (ensemble, aggregate)", "value": round(packed["atom_steps_per_s"], 1),
           "unit": "atom-steps/s", "n_gpus": 1, "steps": args.ens_steps, "warmup": 0,
           "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": dtype_of(args.precision), "data": "synthetic",
           "config": {"workload": f"{len(specs)} independent systems, half charged-LJ salt, half rigid water, "
                                  f"atoms log-uniform in [600, {args.max_atoms}], {args.ens_steps} steps each",
                      "mode": args.ens_mode, "streams": args.streams if args.ens_mode == "streams" else None,
                      "chunk": args.chunk, "precision": args.precision, "plain_loop_precision": "fp64",
                      "hw_queues": int(os.environ.get("GPU_MAX_HW_QUEUES", "4"))},
           "plain_loop_atom_steps_per_s": round(plain["atom_steps_per_s"], 1),
           "speedup_vs_plain_loop": round(packed["atom_steps_per_s"] / plain["atom_steps_per_s"], 3)}
    print(json.dumps(out))


def bench_serve(args):
    """Scaled served stream: Poisson arrivals, admission by footprint,
    continuous batching; p50 / p99 latency and aggregate throughput."""
    from rbmd_amd.serve import build, serve, stream_specs

    jobs = stream_specs(args.rate, args.duration, seed=77, n_max=args.max_atoms, steps_min=200,
                        steps_max=args.max_steps)
    print(f"[bench] serve: {len(jobs)} jobs over {args.duration} s", file=sys.stderr, flush=True)
    inputs = {}
    for j, sp in enumerate(jobs):  # each job's input as it arrives
        inputs[j] = build(sp, args.precision)
        if (j + 1) % 16 == 0:
            print(f"[bench] built {j + 1} jobs", file=sys.stderr, flush=True)
    say = lambda n: print(f"[bench] served {n} jobs", file=sys.stderr, flush=True)
    r = serve(jobs, streams=args.streams, chunk=args.chunk, prebuilt=inputs, precision=args.precision,
              mode=args.ens_mode, progress=say)
    offered = sum(sp.atoms * sp.steps for sp in jobs) / args.duration
    out = {"metric": "served jobs: p50/p99 latency, aggregate atom-steps/s", "value": round(r["atom_steps_per_s"], 1),
           "unit": "atom-steps/s", "n_gpus": 1, "steps": None, "warmup": 0, "ms_per_step": None,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": dtype_of(args.precision),
           "data": "synthetic",
           "config": {"workload": f"Poisson arrivals {args.rate}/s for {args.duration} s, salt or water 50/50, "
                                  f"atoms log-uniform in [1000, {args.max_atoms}], steps uniform in "
                                  f"[200, {args.max_steps}]", "mode": args.ens_mode,
                      "streams": args.streams if args.ens_mode == "streams" else None, "chunk": args.chunk,
                      "precision": args.precision, "hw_queues": int(os.environ.get("GPU_MAX_HW_QUEUES", "4"))},
           "jobs": r["jobs"], "p50_latency_s": r["p50_latency_s"], "p99_latency_s": r["p99_latency_s"],
           "offered_atom_steps_per_s": round(offered, 1),
           "wall_s": round(r["wall_s"], 3)}
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--workload", default="water", choices=["water", "salt", "water300k", "ensemble", "serve"])
    ap.add_argument("--precision", choices=["fp64", "mixed"], default=None,
                    help="default: fp64 for water and salt, mixed for water300k, ensemble and serve")
    ap.add_argument("--systems", type=int, default=64, help="ensemble: number of systems (half salt, half water)")
    ap.add_argument("--ens-steps", type=int, default=200, help="ensemble: MD steps per system")
    ap.add_argument("--ens-mode", choices=["pack", "streams"], default="pack",
                    help="ensemble / serve: one pack per kind (one launch per kernel for all its systems) or "
                         "per-system step graphs spread over --streams contexts")
    ap.add_argument("--streams", type=int, default=8, help="ensemble / serve with --ens-mode streams: contexts")
    ap.add_argument("--chunk", type=int, default=50, help="ensemble / serve: graph steps per enqueue round")
    ap.add_argument("--rate", type=float, default=20.0, help="serve: Poisson arrivals per second")
    ap.add_argument("--duration", type=float, default=5.0, help="serve: seconds of arrivals")
    ap.add_argument("--max-atoms", type=int, default=30000, help="ensemble / serve: largest system")
    ap.add_argument("--max-steps", type=int, default=500, help="serve: most steps of a job")
    ap.add_argument("--cpu-steps", type=int, default=None, help="default: 100 (10 for water300k)")
    ap.add_argument("--rbe-p", type=int, default=None, help="RBE k vectors per step (default: the workload's)")
    ap.add_argument("--hw-queues", type=int, default=None,
                    help="ensemble / serve: hardware queues HIP gives this process (GPU_MAX_HW_QUEUES, 1..32)")
    ap.add_argument("--error-sets", type=int, default=16,
                    help="k sets averaged in the RBE-vs-Ewald force error (0: skip the error report)")
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
    if args.workload == "ensemble":
        return bench_ensemble(args)
    if args.workload == "serve":
        return bench_serve(args)
    system, run, desc, dt_fs = workload(args.workload, args.precision, args.rbe_p)
    if args.rbe_p and args.workload != "water300k":
        desc = desc.replace("RBE P=200", f"RBE P={args.rbe_p}").replace("RBE P=100", f"RBE P={args.rbe_p}")
    print(f"[bench] workload {args.workload}: {system.n_atoms} atoms", file=sys.stderr, flush=True)
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
    print(f"[bench] timed {args.steps} steps in {elapsed:.3f} s", file=sys.stderr, flush=True)

    # per-kernel mean durations with HIP events on the library's stream
    ctx.profile(True)
    ctx.profile_reset()
    g.md_step(args.profile_steps)
    rec = ctx.profile_records()
    ctx.profile(False)
    cands = stencil_candidates(g)
    g.build_neighbors(0)
    entries = g.neighbor_entries()
    kernels = {k: {"launches": n, "mean_us": 1e3 * ms / max(n, 1)} for k, (n, ms) in rec.items()}
    dom = max(rec.items(), key=lambda kv: kv[1][1])[0]
    P = run.rbe_p
    flops = {
        "force_cells": PAIR_FLOPS * entries + TEST_FLOPS * cands,
        "pair": PAIR_FLOPS * entries,
        "nbr_build": TEST_FLOPS * cands,
        "kforce": 35.0 * system.n_atoms * P,
        "rho_partial": 39.0 * system.n_atoms * P,
    }
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    if os.path.exists(tpath) and args.workload == "water" and args.precision == "fp64":  # measured there only
        with open(tpath) as f:
            tt = json.load(f).get("kernels", {}).get(dom)
        if tt:
            traffic = tt["bytes_per_launch"]
    roof = None
    if dom in flops:
        achieved = flops[dom] / (kernels[dom]["mean_us"] * 1e-6) / 1e12
        if args.precision == "mixed" and dom == "pair":
            peak, note = FP32_PEAK_TFLOPS, ("MI355X FP32 vector peak (counts packed FP32, which this build does not "
                                            "emit; the pair scalar is FP32, its vectors and sums FP64)")
        else:
            peak, note = FP64_PEAK_TFLOPS, "MI355X FP64 vector peak (equal to FP64 matrix peak)"
        roof = {"kernel": dom, "bound": "valu", "achieved": round(achieved, 3), "peak": peak,
                "unit": "TFLOP/s", "frac": round(achieved / peak, 4), "traffic": traffic,
                "work_per_launch": flops[dom], "peak_note": note}
    th = g.thermo()
    out = {
        "metric": "atom-steps/s",
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
                   "cut_off": run.cut_off, "tiles": tile_choice(run, system.n_atoms)},
        "ns_per_day": round((args.steps / elapsed) * dt_fs * 86400 / 1e6, 3) if dt_fs else None,
        "roofline": roof,
        "kernels_us": {k: round(v["mean_us"], 2) for k, v in kernels.items()},
        "pairs": entries,
        "stencil_candidates": cands,
        "misfit_atoms": th["misfit_atoms"],
        "final_temperature": th["temperature"],
        "status_flags": th["status_flags"],
    }
    if args.error_sets > 0:
        from rbmd_amd.accuracy import precision_error, rbe_method_error

        # errors on the configuration the timed run ended on
        x, v, _, _ = g.state()
        cur = type(system)(**{**system.__dict__, "x": x, "v": v})
        fe = {}
        if run.coulomb_type == "RBE":
            me = rbe_method_error(ctx, cur, run, n_sets=args.error_sets)
            fe.update({"rbe_vs_ewald_per_step": round(me["per_step_rel_total"], 5),
                       f"rbe_vs_ewald_avg{args.error_sets}": round(me["averaged_rel_total"], 5),
                       "ewald_kmax": me["kmax"]})
        if args.precision == "mixed":
            pe = precision_error(ctx, cur, run)
            fe.update({"mixed_vs_fp64_total": pe.get("total_rel_rms"), "mixed_vs_fp64_real": pe["real_rel_rms"],
                       "mixed_vs_fp64_rbe": pe.get("rbe_rel_rms"), "mixed_vs_fp64_e_lj": pe["e_lj_rel"],
                       "mixed_vs_fp64_e_coul": pe["e_coul_rel"]})
        fe["definition"] = ("relative RMS over atoms against the FP64 reference's nonbonded force "
                            "(real-space LJ+erfc + k space); RBE vs converged FP64 Ewald at the same alpha, "
                            "per step (one k set) and averaged over k sets")
        out["force_error"] = fe
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        cpu_value, cpu_time = cpu_baseline(system, run, args.cpu_steps, threads)
        out["cpu_baseline"] = {"value": round(cpu_value, 1), "unit": "atom-steps/s", "cores": threads,
                               "kind": "port",
                               "sample": f"{args.cpu_steps} MD steps of the same workload after init + 1 step "
                                         f"({cpu_time:.1f} s, OpenMP FP64 oracle)"}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
