"""Many systems on one GPU: every job's result is the same alone and in the
mix, bit for bit: the packed ensemble and the served loop, each with packs
(one launch per kernel for all systems of a kind) and with per-system step
graphs on several streams."""
import numpy as np
import pytest

from rbmd_amd.lib import Context, GpuSystem
from rbmd_amd.serve import JobSpec, build, run_packed, serve

pytestmark = pytest.mark.gpu

JOBS = [JobSpec("salt", 600, 40, 11), JobSpec("water", 603, 30, 12), JobSpec("salt", 2000, 25, 13),
        JobSpec("water", 1500, 35, 14, arrival=0.01)]


def alone(ctx, spec, precision="fp64"):
    s, r = build(spec, precision)
    g = GpuSystem(ctx, s, r)
    g.md_init()
    g.md_step(spec.steps)
    out = g.state()
    g.close()
    return out


@pytest.mark.parametrize("mode", ["pack", "streams"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_packed_equals_alone(ctx, precision, mode):
    print(f"[case] packed ensemble vs alone ({precision}, {mode})", flush=True)
    systems = []
    run_packed(JOBS, streams=2, chunk=7, systems=systems, precision=precision, mode=mode)
    for spec, g in zip(JOBS, systems):
        ref = alone(ctx, spec, precision)
        got = g.state()
        for a, b in zip(ref, got):
            assert np.array_equal(a, b)
    for g in systems:
        g.close()


@pytest.mark.parametrize("mode", ["pack", "streams"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_served_equals_alone(ctx, precision, mode):
    print(f"[case] served stream vs alone ({precision}, {mode})", flush=True)
    results = {}
    r = serve(JOBS, streams=3, chunk=9, results=results, time_scale=1.0, precision=precision, mode=mode)
    assert r["jobs"] == len(JOBS) and r["p99_latency_s"] >= r["p50_latency_s"] > 0
    for j, spec in enumerate(JOBS):
        ref = alone(ctx, spec, precision)
        for a, b in zip(ref, results[j]):
            assert np.array_equal(a, b)
