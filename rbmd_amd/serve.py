"""Many independent systems on one GPU: the plain loop, the packed ensemble and
the served job stream.

Every system owns its device state for its whole life (HBM-resident).  In the
packed and served modes each system is bound to one of K contexts (one HIP
stream each, HIP maps them onto the device's hardware queues), its MD steps
replay one captured step graph (`rbmd_md_enqueue`), and the host only
enqueues: systems on different streams run concurrently, which fills the GPU
with the small kernels of many small systems.  The plain loop is the
reference's way of running them: one system at a time, one stream, steps
enqueued kernel by kernel (`rbmd_md_step`).

The reference has no counterpart for this module (its executioner runs one
system, framework/executioner/Executioner.cpp:62-129); the per-system step is
the same `ExecutionNVT::Solve` sequence as `rbmd_md_step`.
"""
import dataclasses
import math
import time
from typing import List, Optional

import numpy as np

from .config import salt_config, water_config
from .lib import Context, GpuSystem
from .system import make_salt, make_water


@dataclasses.dataclass
class JobSpec:
    kind: str  # "salt" or "water"
    atoms: int
    steps: int
    seed: int
    arrival: float = 0.0  # seconds from the start of the stream


def log_uniform_atoms(rng: np.random.Generator, kind: str, n_min: int, n_max: int) -> int:
    n = int(round(math.exp(rng.uniform(math.log(n_min), math.log(n_max)))))
    if kind == "salt":
        return max(2, n - n % 2)
    return max(3, n - n % 3)


def ensemble_specs(n_systems: int, steps: int, seed: int, n_min: int = 600, n_max: int = 30000) -> List[JobSpec]:
    """The batched-ensemble shape: the first half salt, the second half water,
    atom counts log-uniform in [n_min, n_max] from the seed."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_systems):
        kind = "salt" if i < n_systems // 2 else "water"
        out.append(JobSpec(kind, log_uniform_atoms(rng, kind, n_min, n_max), steps, int(rng.integers(1, 2 ** 31))))
    return out


def stream_specs(rate: float, duration: float, seed: int, n_min: int = 1000, n_max: int = 300000,
                 steps_min: int = 200, steps_max: int = 5000) -> List[JobSpec]:
    """The served-stream shape: Poisson arrivals at `rate` jobs/s for `duration` s,
    salt or water with probability 1/2, atoms log-uniform, steps uniform."""
    rng = np.random.default_rng(seed)
    out, t = [], 0.0
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            return out
        kind = "salt" if rng.random() < 0.5 else "water"
        out.append(JobSpec(kind, log_uniform_atoms(rng, kind, n_min, n_max),
                           int(rng.integers(steps_min, steps_max + 1)), int(rng.integers(1, 2 ** 31)), t))


def build(spec: JobSpec, precision: str = "fp64"):
    if spec.kind == "salt":
        return make_salt(n_ions=spec.atoms, seed=spec.seed), salt_config(precision=precision)
    return make_water(n_mol=spec.atoms // 3, seed=spec.seed), water_config(precision=precision)


def run_plain(specs: List[JobSpec], device: int = 0, precision: str = "fp64") -> dict:
    """One system at a time on one stream, eager steps (rbmd_md_step).  Only
    the MD steps are timed (system set-up and md_init are not)."""
    ctx = Context(device)
    busy = 0.0
    atom_steps = 0
    for sp in specs:
        s, r = build(sp, precision)
        g = GpuSystem(ctx, s, r)
        g.md_init()
        ctx.synchronize()
        t0 = time.perf_counter()
        g.md_step(sp.steps)
        ctx.synchronize()
        busy += time.perf_counter() - t0
        atom_steps += sp.atoms * sp.steps
        g.close()
    ctx.close()
    return {"atom_steps": atom_steps, "seconds": busy, "atom_steps_per_s": atom_steps / busy}


class Pool:
    """K contexts (streams); systems are bound to the least-loaded one."""

    def __init__(self, streams: int, device: int = 0):
        self.ctxs = [Context(device) for _ in range(streams)]
        self.load = [0] * streams

    def pick(self, weight: int) -> int:
        k = int(np.argmin(self.load))
        self.load[k] += weight
        return k

    def release(self, k: int, weight: int):
        self.load[k] -= weight

    def close(self):
        for c in self.ctxs:
            c.close()


def run_packed(specs: List[JobSpec], streams: int = 8, chunk: int = 50, device: int = 0,
               systems: Optional[list] = None, precision: str = "fp64") -> dict:
    """All systems resident at once; every round enqueues `chunk` graph steps
    of every unfinished system on its stream, then waits for the round.  Only
    the stepping is timed."""
    pool = Pool(streams, device)
    live = []
    for sp in specs:
        s, r = build(sp, precision)
        k = pool.pick(sp.atoms)
        g = GpuSystem(pool.ctxs[k], s, r)
        g.md_init()
        live.append([sp, g, 0])
    for c in pool.ctxs:
        c.synchronize()
    t0 = time.perf_counter()
    while True:
        active = [e for e in live if e[2] < e[0].steps]
        if not active:
            break
        for e in active:
            n = min(chunk, e[0].steps - e[2])
            e[1].md_enqueue(n)
            e[2] += n
        for e in active:
            e[1].md_wait()
    seconds = time.perf_counter() - t0
    atom_steps = sum(sp.atoms * sp.steps for sp in specs)
    if systems is not None:
        for e in live:
            e[1]._pool = pool  # the contexts live as long as the returned systems
        systems.extend(e[1] for e in live)
    else:
        for e in live:
            e[1].close()
        pool.close()
    return {"atom_steps": atom_steps, "seconds": seconds, "atom_steps_per_s": atom_steps / seconds,
            "streams": streams, "chunk": chunk}


def serve(jobs: List[JobSpec], streams: int = 8, chunk: int = 50, hbm_budget: float = 200e9, max_live: int = 64,
          device: int = 0, time_scale: float = 1.0, results: Optional[dict] = None,
          prebuilt: Optional[dict] = None, precision: str = "fp64") -> dict:
    """Continuous batching of a job stream.

    Jobs arrive at spec.arrival * time_scale seconds after the start.  A job
    is admitted when a slot (max_live) and HBM (footprint estimate against
    hbm_budget) are free; admitted jobs are stepped `chunk` steps at a time on
    their streams, interleaved with every other live job, and finish after
    their step count.  Latency = finish - arrival.  Returns p50 / p99
    latency and the aggregate atom-steps/s of the run.  `prebuilt` maps a
    job index to its (System, RunConfig), i.e. the job's input as it
    arrives; without it inputs are generated at admission.
    """
    pool = Pool(streams, device)
    pending = sorted(range(len(jobs)), key=lambda i: jobs[i].arrival)
    waiting: List[int] = []
    live = {}
    used = 0.0
    lat = []
    done_atom_steps = 0
    t_start = time.perf_counter()
    nxt = 0
    while nxt < len(pending) or waiting or live:
        now = time.perf_counter() - t_start
        while nxt < len(pending) and jobs[pending[nxt]].arrival * time_scale <= now:
            waiting.append(pending[nxt])
            nxt += 1
        # admission in arrival order while a slot and memory are free
        while waiting and len(live) < max_live:
            j = waiting[0]
            s, r = prebuilt[j] if prebuilt is not None else build(jobs[j], precision)
            need = GpuSystem.estimate_footprint(s, r)
            if live and used + need > hbm_budget:
                break
            waiting.pop(0)
            k = pool.pick(jobs[j].atoms)
            g = GpuSystem(pool.ctxs[k], s, r)
            g.md_init()
            live[j] = [g, 0, k, need]
            used += need
        if not live:
            if nxt < len(pending):
                time.sleep(max(0.0, min(0.01, jobs[pending[nxt]].arrival * time_scale - now)))
            continue
        for j, e in live.items():
            n = min(chunk, jobs[j].steps - e[1])
            e[0].md_enqueue(n)
            e[1] += n
        finished = []
        for j, e in live.items():
            e[0].md_wait()
            if e[1] >= jobs[j].steps:
                finished.append(j)
        for j in finished:
            g, _, k, need = live.pop(j)
            t_done = time.perf_counter() - t_start
            lat.append(t_done - jobs[j].arrival * time_scale)
            done_atom_steps += jobs[j].atoms * jobs[j].steps
            if results is not None:
                results[j] = g.state()
            g.close()
            pool.release(k, jobs[j].atoms)
            used -= need
    wall = time.perf_counter() - t_start
    pool.close()
    lat = np.sort(np.asarray(lat))
    return {"jobs": len(jobs), "wall_s": wall, "atom_steps": done_atom_steps,
            "atom_steps_per_s": done_atom_steps / wall if wall > 0 else 0.0,
            "p50_latency_s": float(np.percentile(lat, 50)) if len(lat) else None,
            "p99_latency_s": float(np.percentile(lat, 99)) if len(lat) else None}
