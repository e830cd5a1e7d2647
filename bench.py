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
HBM_PEAK_GBS = 8000.0
PAIR_FLOPS = 43.0  # per accepted pair (counted from each side): LJ + erfc-Coulomb force
TEST_FLOPS = 8.0  # per distance test of an atom against a stencil candidate


def workload(name):
    from rbmd_amd import make_salt, make_water, salt_config, water_config

    if name == "water":
        s = make_water(n_mol=10000, seed=1)
        r = water_config()
        desc = ("rigid 3-site water, 10,000 molecules (30,000 atoms), 1.0 g/cm3 (L=66.88 A), rc 10 A, "
                "alpha 0.09 (beta 0.3/A), RBE P=200, SHAKE, NVT Nose-Hoover 300 K, dt 2 fs")
        dt_fs = 2.0
    elif name == "salt":
        s = make_salt(n_ions=4000, seed=1)
        r = salt_config()
        desc = ("charged-LJ salt, 4,000 ions, rho*=0.5 (L=20 sigma), rc 4 sigma, alpha 0.64 (beta 0.8), "
                "RBE P=100, NVT Nose-Hoover T*=1.5, dt 0.001 tau")
        dt_fs = None
    else:
        raise SystemExit(f"unknown workload {name}")
    return s, r, desc, dt_fs


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--workload", default="water", choices=["water", "salt"])
    ap.add_argument("--cpu-steps", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-steps", type=int, default=50)
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

    from rbmd_amd.lib import Context, GpuSystem

    system, run, desc, dt_fs = workload(args.workload)
    print(f"[bench] workload {args.workload}: {system.n_atoms} atoms", file=sys.stderr, flush=True)
    ctx = Context(0)
    g = GpuSystem(ctx, system, run)
    g.md_init()
    g.md_step(args.warmup)
    ctx.synchronize()
    t0 = time.perf_counter()
    g.md_step(args.steps)
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
    if os.path.exists(tpath):
        with open(tpath) as f:
            tt = json.load(f).get("kernels", {}).get(dom)
        if tt:
            traffic = tt["bytes_per_launch"]
    roof = None
    if dom in flops:
        achieved = flops[dom] / (kernels[dom]["mean_us"] * 1e-6) / 1e12
        roof = {"kernel": dom, "bound": "valu", "achieved": round(achieved, 3), "peak": FP64_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": round(achieved / FP64_PEAK_TFLOPS, 4), "traffic": traffic,
                "work_per_launch": flops[dom], "peak_note": "MI355X FP64 vector peak (equal to FP64 matrix peak)"}
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
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": desc, "atoms": system.n_atoms, "precision": "fp64", "rbe_p": P,
                   "cut_off": run.cut_off},
        "ns_per_day": round((args.steps / elapsed) * dt_fs * 86400 / 1e6, 3) if dt_fs else None,
        "roofline": roof,
        "kernels_us": {k: round(v["mean_us"], 2) for k, v in kernels.items()},
        "pairs": entries,
        "stencil_candidates": cands,
        "misfit_atoms": th["misfit_atoms"],
        "final_temperature": th["temperature"],
        "status_flags": th["status_flags"],
    }
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
