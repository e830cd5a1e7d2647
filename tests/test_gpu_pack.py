"""Packs (rbmd_pack_*): many systems stepped with one launch per kernel of
the step.  Every system's state and thermo after packed steps must equal the
same system stepped alone, bit for bit: eager and graph-replayed packs, FP64
and mixed precision, systems of different sizes (including a box below 2 rc),
systems that join at different step counts, and runs that cross the k-vector
ring refill."""
import numpy as np
import pytest

from rbmd_amd.lib import GpuSystem, Pack, RbmdError
from tests.systems import salt, salt_run, water, water_run

pytestmark = pytest.mark.gpu


def alone(ctx, system, run, steps, graph=False):
    g = GpuSystem(ctx, system, run)
    g.md_init()
    (g.md_run_graph if graph else g.md_step)(steps)
    out = g.state(), g.thermo()
    g.close()
    return out


def same(ref, g):
    (rx, rv, rf, ri), rt = ref
    gx, gv, gf, gi = g.state()
    gt = g.thermo()
    for a, b in ((rx, gx), (rv, gv), (rf, gf), (ri, gi)):
        assert np.array_equal(a, b)
    for k in ("step", "temperature", "pressure", "e_lj", "e_coul", "e_bond", "e_angle"):
        assert rt[k] == gt[k], k
    assert list(rt["nhc_eta"]) == list(gt["nhc_eta"])


def make(ctx, specs, run_of):
    gs = []
    for s in specs:
        g = GpuSystem(ctx, s, run_of())
        g.md_init()
        gs.append(g)
    return gs


@pytest.mark.parametrize("graph", [False, True])
def test_pack_salt_equals_alone(ctx, graph):
    specs = [salt(n=600, seed=21), salt(n=1000, seed=22), salt(n=2000, seed=23)]
    print(f"[case] salt pack graph={graph}", flush=True)
    gs = make(ctx, specs, salt_run)
    p = Pack(ctx, gs)
    if graph:
        p.enqueue(30)
        p.wait()
    else:
        p.step(30)
    for s, g in zip(specs, gs):
        same(alone(ctx, s, salt_run(), 30), g)
    p.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_pack_water_equals_alone(ctx, precision):
    # 201 molecules: L = 18.2 A < 2 rc, one cell per axis, minimum-image pair vectors
    specs = [water(nmol=201, seed=8), water(nmol=300, seed=31), water(nmol=600, seed=5)]
    run = lambda: water_run(precision=precision)
    print(f"[case] water pack {precision}", flush=True)
    gs = make(ctx, specs, run)
    p = Pack(ctx, gs)
    p.enqueue(12)
    p.wait()
    for s, g in zip(specs, gs):
        same(alone(ctx, s, run(), 12, graph=True), g)


def test_pack_staggered_steps_and_ring_refill(ctx):
    """systems join at different step counts; 300 packed steps cross the
    256-step k-vector ring of both"""
    a, b = salt(n=600, seed=41), salt(n=800, seed=42)
    print("[case] staggered salt pack", flush=True)
    ga = GpuSystem(ctx, a, salt_run())
    ga.md_init()
    ga.md_step(7)
    gb = GpuSystem(ctx, b, salt_run())
    gb.md_init()
    p = Pack(ctx, [ga, gb])
    p.enqueue(150)
    p.wait()
    p.step(150)
    same(alone(ctx, a, salt_run(), 307), ga)
    same(alone(ctx, b, salt_run(), 300), gb)


def test_pack_rejects_mixed_shapes(ctx):
    ga = GpuSystem(ctx, salt(n=600, seed=21), salt_run())
    gw = GpuSystem(ctx, water(nmol=300, seed=31), water_run())
    with pytest.raises(RbmdError):
        Pack(ctx, [ga])  # not initialised
    ga.md_init()
    gw.md_init()
    with pytest.raises(RbmdError):
        Pack(ctx, [ga, gw])
    with pytest.raises(RbmdError):
        Pack(ctx, [ga, ga])


def test_tile_options_agree(ctx):
    """every tile option gives the same forces up to summation order"""
    s = water(nmol=600, seed=5)
    ref = None
    for i, (pl, kl, kb) in enumerate([(4, 2, 1), (8, 4, 2), (16, 8, 4)]):
        print(f"[case] tiles pair {pl} kforce {kl} rho {kb}", flush=True)
        g = GpuSystem(ctx, s, water_run(pair_lanes=pl, kforce_lanes=kl, rho_kblock=kb))
        g.md_init()
        f = g.state()[2]
        g.close()
        if ref is None:
            ref = f
        else:
            assert np.abs(f - ref).max() <= 1e-12 * np.abs(ref).max()


def test_pack_rejects_different_k_tiles(ctx):
    a = GpuSystem(ctx, salt(n=600, seed=21), salt_run(kforce_lanes=2))
    b = GpuSystem(ctx, salt(n=800, seed=22), salt_run(kforce_lanes=8))
    a.md_init()
    b.md_init()
    with pytest.raises(RbmdError):
        Pack(ctx, [a, b])
