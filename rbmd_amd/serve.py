"""Many independent systems on one GPU: the plain loop, the packed ensemble and
the served job stream.

Every system owns its device state for its whole life (HBM-resident).  In
the packed and served modes the systems of one kind (salt or water: the same
force field, constraints and thermostat) form a pack (rbmd_pack_*): every
kernel of the MD step is one launch over the blocks of all of them, replayed
from one captured graph, so a step of many small systems costs one chain of
launches instead of one chain each.  Each kind's packs have their own context
(HIP stream); rounds run one pack at a time (step_group; the served loop picks
each round's pack by remaining work).  The older mode
("streams") binds each system to one of K contexts and replays its own step
graph; it is kept for comparison.  The plain loop is the reference's way of
running many systems: one at a time, one stream, steps enqueued kernel by
kernel (rbmd_md_step).

The reference has no counterpart for this module (its executioner runs one
system, framework/executioner/Executioner.cpp:62-129); each system's step is
the same ExecutionNVT::Solve sequence as rbmd_md_step, with the same bits.
"""
import collections
import dataclasses
import math
import time
from typing import List, Optional

import numpy as np

from .config import salt_config, water_config
from .lib import Context, GpuSystem, Pack
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


def water_rbe_p(atoms: int) -> int:
    """RBE k vectors per step for a water job: 200 up to 30,000 atoms, then
    growing with N.  At fixed P the per-step RBE force error grows like
    sqrt(N) (0.34 / 1.09 / 3.3 of the force RMS at 3k / 30k / 300k atoms, P =
    200); at 300k atoms and P = 200 the lattice start fails SHAKE within 10
    steps, in the CPU oracle as on the GPU, while P = 200 N / 30,000 keeps the
    30,000-atom error level."""
    return max(200, int(round(200 * atoms / 30000)))


def build(spec: JobSpec, precision: str = "fp64"):
    if spec.kind == "salt":
        return make_salt(n_ions=spec.atoms, seed=spec.seed), salt_config(precision=precision)
    return (make_water(n_mol=spec.atoms // 3, seed=spec.seed),
            water_config(precision=precision, rbe_p=water_rbe_p(spec.atoms)))


def build_all(specs: List[JobSpec], precision: str = "fp64", progress=None) -> list:
    """every job's (System, RunConfig); progress(i) is called every 32 builds"""
    out = []
    for i, sp in enumerate(specs):
        out.append(build(sp, precision))
        if progress is not None and (i + 1) % 32 == 0:
            progress(i + 1)
    return out


def run_plain(specs: List[JobSpec], device: int = 0, precision: str = "fp64", inputs: Optional[list] = None,
              progress=None) -> dict:
    """One system at a time on one stream, eager steps (rbmd_md_step).  Only
    the MD steps are timed (system set-up and md_init are not).  inputs: the
    jobs' (System, RunConfig) (built here when absent)."""
    ctx = Context(device)
    busy = 0.0
    atom_steps = 0
    for i, sp in enumerate(specs):
        s, r = inputs[i] if inputs is not None else build(sp, precision)
        g = GpuSystem(ctx, s, r)
        g.md_init()
        ctx.synchronize()
        t0 = time.perf_counter()
        g.md_step(sp.steps)
        ctx.synchronize()
        busy += time.perf_counter() - t0
        atom_steps += sp.atoms * sp.steps
        g.close()
        if progress is not None and (i + 1) % 64 == 0:
            progress(i + 1)
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


class PackGroup:
    """The live systems of one kind on one context, stepped as one pack.  Each
    round steps the pack by min(chunk, the smallest remaining step count), so
    every member stops exactly at its own count; the pack is rebuilt whenever
    members join or leave."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        self.members = []  # [key, GpuSystem, remaining steps]
        self.pack = None
        self.n = 0

    def add(self, key, g, steps):
        self.members.append([key, g, steps])
        self._drop_pack()

    def _drop_pack(self):
        if self.pack is not None:
            self.pack.close()
            self.pack = None

    def launch(self, chunk):
        self.n = 0
        if not self.members:
            return
        if self.pack is None:
            self.pack = Pack(self.ctx, [m[1] for m in self.members])
        self.n = min(chunk, min(m[2] for m in self.members))
        self.pack.enqueue(self.n)

    def finish(self):
        """wait for the round; returns the members that reached their count"""
        if not self.n:
            return []
        self.pack.wait()
        done = []
        for m in self.members:
            m[2] -= self.n
            if m[2] <= 0:
                done.append(m)
        if done:
            self.members = [m for m in self.members if m[2] > 0]
            self._drop_pack()
        return done

    def close(self):
        self._drop_pack()


def step_group(gr: "PackGroup", chunk: int) -> list:
    """One round of one kind's pack: launch it and wait for it before the next
    kind's round is launched, so at most one pack has work in flight.  Two
    256-system packs in flight together (one per kind, each on its own
    context) stopped making progress once in a full benchmark process, and
    no cause was found in this code; rounds of one pack at a time are the
    schedule the full-size ensemble test runs."""
    gr.launch(chunk)
    return gr.finish()


def run_packed(specs: List[JobSpec], streams: int = 8, chunk: int = 50, device: int = 0,
               systems: Optional[list] = None, precision: str = "fp64", mode: str = "pack",
               inputs: Optional[list] = None, progress=None) -> dict:
    """All systems resident at once.  mode "pack": one pack per kind, each on
    its own context, `chunk` graph-replayed packed steps per round; mode
    "streams": systems spread over `streams` contexts, `chunk` steps of each
    system's own step graph per round.  Only the stepping is timed."""
    if mode not in ("pack", "streams"):
        raise ValueError(mode)
    built = []
    if mode == "pack":
        kinds = sorted({sp.kind for sp in specs})
        ctxs = {k: Context(device) for k in kinds}
        groups = {k: PackGroup(ctxs[k]) for k in kinds}
        for i, sp in enumerate(specs):
            s, r = inputs[i] if inputs is not None else build(sp, precision)
            g = GpuSystem(ctxs[sp.kind], s, r)
            g.md_init()
            built.append(g)
            groups[sp.kind].add(i, g, sp.steps)
            if progress is not None and (i + 1) % 128 == 0:
                progress(f"{i + 1} systems set up")
        for c in ctxs.values():
            c.synchronize()
        t0 = time.perf_counter()
        rounds = 0
        while any(gr.members for gr in groups.values()):
            for gr in groups.values():  # one kind's round at a time (see step_group)
                step_group(gr, chunk)
            rounds += 1
            if progress is not None and rounds % 2 == 0:
                progress(f"{rounds} rounds of {chunk} steps")
        seconds = time.perf_counter() - t0
        for gr in groups.values():
            gr.close()
        holders = list(ctxs.values())
    else:
        pool = Pool(streams, device)
        live = []
        for i, sp in enumerate(specs):
            s, r = inputs[i] if inputs is not None else build(sp, precision)
            k = pool.pick(sp.atoms)
            g = GpuSystem(pool.ctxs[k], s, r)
            g.md_init()
            built.append(g)
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
        holders = [pool]
    atom_steps = sum(sp.atoms * sp.steps for sp in specs)
    if systems is not None:
        for g in built:
            g._holders = holders  # the contexts live as long as the returned systems
        systems.extend(built)
    else:
        for g in built:
            g.close()
        for h in holders:
            h.close()
    return {"atom_steps": atom_steps, "seconds": seconds, "atom_steps_per_s": atom_steps / seconds,
            "mode": mode, "contexts": len(holders) if mode == "pack" else streams, "chunk": chunk}


def serve(jobs: List[JobSpec], streams: int = 8, chunk: int = 50, hbm_budget: float = 200e9, max_live: int = 64,
          device: int = 0, time_scale: float = 1.0, results: Optional[dict] = None,
          prebuilt: Optional[dict] = None, precision: str = "fp64", mode: str = "pack", progress=None,
          small_atoms: int = 20000, max_packs: int = 16, age_rate: float = 1.5e8) -> dict:
    """Continuous batching of a job stream.

    Jobs arrive at spec.arrival * time_scale seconds after the start.  A job
    is admitted when a slot (max_live) and HBM (footprint estimate against
    hbm_budget) are free, and finishes after its step count.  Mode "pack":
    each round runs, for up to `chunk` steps, the live job with the least
    remaining work (atoms x steps left, less age_rate atom-steps per second
    it has waited, so a large job is not starved by a stream of small ones),
    in one pack with the live jobs of its kind that are at most twice its
    size (or up to `small_atoms`): a pack steps about as fast as its
    smallest member, and small jobs are not held to a large job's step time.
    Packs are kept while their members are live; a job alone steps with its
    own launches.  Mode "streams": every live job is stepped `chunk` steps per round
    on one of `streams` contexts with its own step graph.  Latency = finish -
    arrival.  Returns p50 / p99 latency and the aggregate atom-steps/s of the
    run.  `prebuilt` maps a job index to its (System, RunConfig), i.e. the
    job's input as it arrives; without it inputs are generated at admission.
    """
    if mode not in ("pack", "streams"):
        raise ValueError(mode)
    pending = sorted(range(len(jobs)), key=lambda i: jobs[i].arrival)
    waiting: List[int] = []
    live = {}  # job -> [GpuSystem, steps done, context slot, footprint]
    used = 0.0
    lat = []
    done_atom_steps = 0
    packs = collections.OrderedDict()  # (kind, member jobs) -> Pack, while every member is live
    if mode == "pack":
        kinds = sorted({j.kind for j in jobs})
        ctxs = {k: Context(device) for k in kinds}
        pool = None
    else:
        pool = Pool(streams, device)
    t_start = time.perf_counter()
    nxt = 0

    def finish_job(j):
        nonlocal used, done_atom_steps
        for key in [key for key in packs if j in key[1]]:  # a pack must not outlive its systems
            packs.pop(key).close()
        g, _, k, need = live.pop(j)
        lat.append((time.perf_counter() - t_start) - jobs[j].arrival * time_scale)
        done_atom_steps += jobs[j].atoms * jobs[j].steps
        if results is not None:
            results[j] = g.state()
        g.close()
        if pool is not None:
            pool.release(k, jobs[j].atoms)
        used -= need
        if progress is not None and len(lat) % 10 == 0:
            progress(len(lat))

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
            if pool is not None:
                k = pool.pick(jobs[j].atoms)
                g = GpuSystem(pool.ctxs[k], s, r)
            else:
                k = jobs[j].kind
                g = GpuSystem(ctxs[k], s, r)
            g.md_init()
            live[j] = [g, 0, k, need]
            used += need
        if not live:
            if nxt < len(pending):
                time.sleep(max(0.0, min(0.01, jobs[pending[nxt]].arrival * time_scale - now)))
            continue
        if pool is None:
            t_now = time.perf_counter() - t_start

            def urgency(j):  # remaining work, less what the wait so far is worth
                left = (jobs[j].steps - live[j][1]) * jobs[j].atoms
                return (left - age_rate * (t_now - jobs[j].arrival * time_scale), jobs[j].arrival)

            head = min(live, key=urgency)
            kind = jobs[head].kind
            limit = max(2 * jobs[head].atoms, small_atoms)
            members = tuple(sorted(j for j in live if jobs[j].kind == kind and jobs[j].atoms <= limit))
            n = min(chunk, min(jobs[j].steps - live[j][1] for j in members))
            if len(members) == 1:  # one system: its own launches (a replayed graph is no faster there)
                live[head][0].md_step(n)
                ctxs[kind].synchronize()
            else:
                key = (kind, members)
                if key in packs:
                    packs.move_to_end(key)
                else:
                    packs[key] = Pack(ctxs[kind], [live[j][0] for j in members])
                    while len(packs) > max_packs:  # the least recently used pack goes
                        packs.popitem(last=False)[1].close()
                packs[key].enqueue(n)
                packs[key].wait()
            for j in members:
                live[j][1] += n
            for j in members:
                if live[j][1] >= jobs[j].steps:
                    finish_job(j)
        else:
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
                finish_job(j)
    wall = time.perf_counter() - t_start
    for pk in packs.values():
        pk.close()
    if pool is not None:
        pool.close()
    else:
        for c in ctxs.values():
            c.close()
    lat = np.sort(np.asarray(lat))
    return {"jobs": len(jobs), "wall_s": wall, "atom_steps": done_atom_steps,
            "atom_steps_per_s": done_atom_steps / wall if wall > 0 else 0.0,
            "p50_latency_s": float(np.percentile(lat, 50)) if len(lat) else None,
            "p99_latency_s": float(np.percentile(lat, 99)) if len(lat) else None, "mode": mode}


def serve_plain(jobs: List[JobSpec], device: int = 0, time_scale: float = 1.0, prebuilt: Optional[dict] = None,
                results: Optional[dict] = None, progress=None) -> dict:
    """The same job stream run the plain way: one system at a time on one
    stream, in arrival order, FP64, eager steps (rbmd_md_step), each job set up
    when the previous one has finished.  Arrivals follow the same clock as
    serve(); latency = finish - arrival, so jobs that arrive while another runs
    wait in the queue.  `prebuilt` maps a job index to its (System, RunConfig)
    (the precision is forced to FP64)."""
    order = sorted(range(len(jobs)), key=lambda i: jobs[i].arrival)
    ctx = Context(device)
    lat = []
    busy = 0.0
    done_atom_steps = 0
    t_start = time.perf_counter()
    for j in order:
        now = time.perf_counter() - t_start
        wait = jobs[j].arrival * time_scale - now
        if wait > 0:
            time.sleep(wait)
        s, r = prebuilt[j] if prebuilt is not None else build(jobs[j], "fp64")
        if r.precision != "fp64":
            r = type(r)(**{**r.__dict__, "precision": "fp64"})
        t0 = time.perf_counter()
        g = GpuSystem(ctx, s, r)
        g.md_init()
        g.md_step(jobs[j].steps)
        ctx.synchronize()
        if results is not None:
            results[j] = g.state()
        g.close()
        t1 = time.perf_counter()
        busy += t1 - t0
        lat.append((t1 - t_start) - jobs[j].arrival * time_scale)
        done_atom_steps += jobs[j].atoms * jobs[j].steps
        if progress is not None and len(lat) % 10 == 0:
            progress(len(lat))
    wall = time.perf_counter() - t_start
    ctx.close()
    lat = np.sort(np.asarray(lat))
    return {"jobs": len(jobs), "wall_s": wall, "busy_s": busy, "atom_steps": done_atom_steps,
            "atom_steps_per_s": done_atom_steps / wall if wall > 0 else 0.0,
            "p50_latency_s": float(np.percentile(lat, 50)) if len(lat) else None,
            "p99_latency_s": float(np.percentile(lat, 99)) if len(lat) else None, "mode": "plain"}


def _build_one(args):
    spec, precision = args
    return build(spec, precision)


def build_parallel(specs: List[JobSpec], precision: str = "fp64", workers: int = 16, progress=None) -> list:
    """build_all over a pool of fresh interpreters (spawned, so no GPU state is
    inherited); the same inputs as build_all, in order."""
    import concurrent.futures as cf
    import multiprocessing as mp

    workers = max(1, min(workers, len(specs)))
    if workers == 1:
        return build_all(specs, precision, progress)
    out = []
    with cf.ProcessPoolExecutor(max_workers=workers, mp_context=mp.get_context("spawn")) as ex:
        for i, r in enumerate(ex.map(_build_one, [(sp, precision) for sp in specs], chunksize=4)):
            out.append(r)
            if progress is not None and (i + 1) % 64 == 0:
                progress(i + 1)
    return out
