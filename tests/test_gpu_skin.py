"""Verlet lists with a skin (params.skin > 0) against the oracle's per-step
lists.

The skin list holds every pair within rc + skin and is rebuilt on the device
when an atom has moved more than skin / 2 since the last build; the pair
kernel applies the reference's test rc^2 - d^2 > 1e-4 to every entry, so each
step sums exactly the pairs of the reference's per-step list (in a different
order).  Bars: forces within 1e-10 after md_init, MD trajectories within
1e-8 after 20 steps (as the per-step path), graph replays bit-identical to
eager steps, far fewer list builds than steps, and the parity exports
(locator, lists, force_real) still give the reference's per-step results
after skin steps.
"""
import numpy as np
import pytest

import oracle
from rbmd_amd.lib import GpuSystem, RbmdError
from tests.systems import rel_err, salt, salt_run, water, water_run
from tests.test_gpu_parity import _md_compare, pair

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind", ["salt", "water"])
def test_skin_md_matches_oracle(ctx, kind):
    s, r = (salt(), salt_run(skin=0.5)) if kind == "salt" else (water(), water_run(skin=1.0))
    t = _md_compare(ctx, s, r, 20, 1e-8)
    assert t["status_flags"] == 0
    assert 1 <= t["list_builds"] < 21  # md_init's build, then rebuilds only when needed


def test_skin_rebuilds_on_displacement(ctx):
    """over 300 water steps the list is rebuilt when atoms have moved, not every
    step (this box relaxes from its lattice start to ~950 K, where the fastest
    hydrogen moves ~0.25 A per step: a 2 A skin lasts a few steps); then, on a
    step without a rebuild, the step's own force (summed over a list kept for
    several steps) equals the force of a list built on the same positions.
    Ewald k space makes the force a function of the positions alone.  The two
    paths' trajectories are not compared this far out: their summation-order
    differences grow exponentially in a chaotic system (the 20-step comparison
    above bounds them while they are small)."""
    s = water()
    r = water_run(skin=2.0, coulomb_type="EWALD", kmax=[6, 6, 6])
    g1 = GpuSystem(ctx, s, r)
    print("[case] skin water 300 steps, Ewald", flush=True)
    g1.md_init()
    for _ in range(3):
        g1.md_step(100)
    t1 = g1.thermo()
    assert 10 <= t1["list_builds"] <= 150, t1["list_builds"]
    assert t1["status_flags"] == 0
    kept = False
    for _ in range(20):  # a step that kept the list
        before = g1.thermo()["list_builds"]
        g1.md_step(1)
        if g1.thermo()["list_builds"] == before:
            kept = True
            break
    assert kept
    x1, v1, f1, _ = g1.state()
    cur = type(s)(**{**s.__dict__, "x": x1, "v": v1})
    g2 = GpuSystem(ctx, cur, water_run(coulomb_type="EWALD", kmax=[6, 6, 6]))
    g2.md_init()  # the force of a list built on these positions
    _, _, f2, _ = g2.state()
    assert rel_err(f1, f2) < 1e-10


@pytest.mark.parametrize("kind", ["salt", "water"])
def test_skin_graph_matches_eager(ctx, kind):
    s, r = (salt(n=2000, seed=4), salt_run(skin=0.5)) if kind == "salt" else (water(), water_run(skin=1.0))
    print(f"[case] skin graph vs eager {kind}", flush=True)
    g1, g2 = GpuSystem(ctx, s, r), GpuSystem(ctx, s, r)
    g1.md_init()
    g2.md_init()
    for n in (7, 13, 30):
        g1.md_step(n)
        g2.md_run_graph(n)
        for a, b in zip(g1.state(), g2.state()):
            assert np.array_equal(a, b)
        t1, t2 = g1.thermo(), g2.thermo()
        for k in ("step", "temperature", "pressure", "e_lj", "e_coul", "list_builds"):
            assert t1[k] == t2[k], k


def test_skin_mixed_precision(ctx):
    s = water()
    gm = GpuSystem(ctx, s, water_run(skin=1.0, precision="mixed"))
    gf = GpuSystem(ctx, s, water_run(skin=1.0))
    gm.md_init()
    gf.md_init()
    _, _, fm, _ = gm.state()
    _, _, ff, _ = gf.state()
    nb = np.linalg.norm(ff, axis=1)
    assert np.sqrt((np.linalg.norm(fm - ff, axis=1) ** 2).mean() / (nb ** 2).mean()) < 1e-5
    gm.md_step(20)
    assert gm.thermo()["status_flags"] == 0


def test_skin_exports_stay_per_step(ctx):
    """after skin steps the locator, list and real-space exports are the
    reference's per-step results on the current state; MD then continues"""
    s, r = salt(), salt_run(skin=0.5)
    g = GpuSystem(ctx, s, r)
    g.md_init()
    g.md_step(15)
    x, v, _, _ = g.state()
    cur = type(s)(**{**s.__dict__, "x": x, "v": v})
    o = oracle.OracleSystem(cur, salt_run())
    assert g.geometry()[0].tolist() == o.geometry()[0].tolist() == [5, 5, 5]
    for a, b in zip(g.locator(), o.locator()):
        assert np.array_equal(a, b)
    gi, go, gn = g.neighbors(400)
    oi, oo, on = o.neighbors(400)
    assert np.array_equal(gn, on) and np.array_equal(gi, oi) and np.array_equal(go, oo)
    gf, gv, _, _ = g.force_real()
    of, ov, _, _ = o.force_real()
    assert rel_err(gf, of) < 1e-10
    g.md_step(5)
    assert g.thermo()["status_flags"] == 0 and g.thermo()["step"] == 20


def test_skin_refused_where_not_built(ctx):
    from rbmd_amd.abi import RBMD_ERR_UNSUPPORTED

    with pytest.raises(RbmdError) as e:  # box 18.2 A < 2 (rc + skin)
        GpuSystem(ctx, water(nmol=201, seed=8), water_run(skin=1.0))
    assert e.value.code == RBMD_ERR_UNSUPPORTED
    with pytest.raises(RbmdError) as e:
        GpuSystem(ctx, salt(), salt_run(skin=0.5, pair_path="cells"))
    assert e.value.code == RBMD_ERR_UNSUPPORTED
