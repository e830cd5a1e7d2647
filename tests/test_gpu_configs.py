"""Every workload at its stated size on the GPU.

  water 30k FP64 (10,000 molecules, L = 66.88 A, rc 10, P = 200, SHAKE)
      locator, Verlet lists (ids, order, images), real-space force, RBE force on
      an injected k set and 5 NVT steps against the oracle; SHAKE at the
      tolerance the workload names (1e-8)
  water 300k mixed (100,000 molecules, L = 144.09 A, P = 2000)
      locator (2,744 cells: the separate scan and scatter kernels) and one
      FP64 force evaluation against the oracle; mixed vs FP64 forces within
      1e-5 relative RMS; momentum; SHAKE bond residuals after mixed steps
  salt 64,000 ions at rc 2.5 (8,000 cells: the atomic scatter without LDS
      aggregation) locator, lists and forces against the oracle
  512-system ensemble (256 salt + 256 water, atoms log-uniform in [600,
      30,000], mixed, 500 steps, packed graph steps): the smallest member of
      each kind against the oracle after 10 steps, and a sample of members bit
      for bit against the same system stepped alone after all 500
  served stream (arrivals, sizes and step counts of the served workload, the
      rate scaled from 20 to 0.5 jobs/s): every job served, p50/p99 reported,
      and the smallest job of each kind bit for bit against running alone
"""
import numpy as np
import pytest

import oracle
from rbmd_amd import make_salt, make_water, salt_config, water_config
from rbmd_amd.lib import GpuSystem
from tests.systems import rel_err
from tests.test_gpu_parity import _md_compare

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10


def _threads():
    oracle.set_threads(16)


def _rel_rms(a, b):
    return float(np.sqrt(((a - b) ** 2).sum(1).mean() / (b ** 2).sum(1).mean()))


def _bond_residual(s, x):
    ang = s.angle_atoms.reshape(-1, 3)  # H, O, H
    box = s.box

    def vec(a, b):
        d = x[b] - x[a]
        return d - box * np.round(d / box)

    r1 = np.linalg.norm(vec(ang[:, 1], ang[:, 0]), axis=1)
    r2 = np.linalg.norm(vec(ang[:, 1], ang[:, 2]), axis=1)
    return max(np.abs(r1 - 1.0).max(), np.abs(r2 - 1.0).max())


@pytest.fixture(scope="module")
def water30k():
    return make_water(n_mol=10000, seed=1)


@pytest.mark.timeout(600)
def test_water30k_lists_forces(ctx, water30k):
    _threads()
    s, r = water30k, water_config()
    print(f"[case] water 30k locator/lists/forces n={s.n_atoms}", flush=True)
    g, o = GpuSystem(ctx, s, r), oracle.OracleSystem(s, r)
    assert g.geometry()[0].tolist() == [6, 6, 6]
    for a, b in zip(g.locator(), o.locator()):
        assert np.array_equal(a, b)
    cap = 600
    gi, go, gn = g.neighbors(cap)
    oi, oo, on = o.neighbors(cap)
    assert np.array_equal(gn, on)
    assert np.array_equal(gi, oi)
    assert np.array_equal(go, oo)
    del gi, go, oi, oo
    gf, gv, ge, gc = g.force_real()
    of, ov, oe, oc = o.force_real()
    assert rel_err(gf, of) < FP64_TOL
    assert rel_err(gv, ov) < FP64_TOL
    assert abs(ge.sum() - oe.sum()) <= FP64_TOL * np.abs(oe).sum()
    assert abs(gc.sum() - oc.sum()) <= FP64_TOL * np.abs(oc).sum()
    m = oracle.rbe_sample(r.alpha, s.box, r.rbe_p, 31, 7)
    g.set_kvectors(m)
    o.set_kvectors(m)
    kf, kv = g.rbe_force()
    of, ov = o.rbe_force()
    assert rel_err(kf, of) < FP64_TOL
    assert rel_err(kv, ov) < 1e-9


@pytest.mark.timeout(600)
def test_water30k_nvt_steps(ctx, water30k):
    _threads()
    print("[case] water 30k 5 NVT steps vs oracle", flush=True)
    t = _md_compare(ctx, water30k, water_config(), 5, 1e-8)
    assert t["status_flags"] == 0


@pytest.mark.timeout(600)
def test_water30k_shake_tolerance(ctx, water30k):
    """SHAKE at the workload's tolerance 1e-8 on the multiplier change: bond
    residuals fall accordingly (the reference's fixed 1e-5 leaves ~1e-6)"""
    print("[case] water 30k SHAKE tolerance 1e-8", flush=True)
    res = {}
    for tol in (0.0, 1e-8):
        g = GpuSystem(ctx, water30k, water_config(shake_tol=tol))
        g.md_init()
        g.md_step(10)
        x, _, _, _ = g.state()
        res[tol] = _bond_residual(water30k, x)
        assert g.thermo()["status_flags"] == 0
        g.close()
    assert res[1e-8] < 1e-9
    assert res[1e-8] < 0.01 * res[0.0]


@pytest.mark.timeout(900)
def test_water300k_mixed(ctx):
    _threads()
    s = make_water(n_mol=100000, seed=1)
    print(f"[case] water 300k mixed n={s.n_atoms}", flush=True)
    rf, rm = water_config(rbe_p=2000), water_config(rbe_p=2000, precision="mixed")
    gf = GpuSystem(ctx, s, rf)
    assert gf.geometry()[0].tolist() == [14, 14, 14]
    o = oracle.OracleSystem(s, rf)
    for a, b in zip(gf.locator(), o.locator()):
        assert np.array_equal(a, b)
    # one FP64 force evaluation against the oracle (real space, RBE on a P = 200 set)
    o.L.orc_neighbors_build(o.h, 600)
    gf.build_neighbors(0)
    f64, v64, _, _ = gf.force_real()
    of, ov, _, _ = o.force_real()
    assert rel_err(f64, of) < FP64_TOL
    assert rel_err(v64, ov) < FP64_TOL
    del o
    m = oracle.rbe_sample(rf.alpha, s.box, 200, 41, 3)
    o2 = oracle.OracleSystem(s, rf)
    o2.set_kvectors(m)
    gf.set_kvectors(m)
    k64, _ = gf.rbe_force()
    ok, _ = o2.rbe_force()
    assert rel_err(k64, ok) < FP64_TOL
    del o2
    # mixed against FP64 at full size: real space, RBE at P = 2000, momentum
    gm = GpuSystem(ctx, s, rm)
    gm.build_neighbors(0)
    f32, _, _, _ = gm.force_real()
    assert _rel_rms(f32, f64) < 1e-5
    m2 = oracle.rbe_sample(rf.alpha, s.box, 2000, 42, 5)
    gm.set_kvectors(m2)
    gf.set_kvectors(m2)
    k32, _ = gm.rbe_force()
    k64, _ = gf.rbe_force()
    assert _rel_rms(k32, k64) < 1e-5
    assert _rel_rms(f32 + k32, f64 + k64) < 1e-5
    scale = np.abs(f64).sum()
    assert np.abs(f32.sum(0)).max() < 1e-9 * scale and np.abs(k32.sum(0)).max() < 1e-9 * np.abs(k64).sum()
    gf.close()
    gm.close()
    # mixed MD: constraints hold and nothing is flagged
    gm = GpuSystem(ctx, s, rm)
    gm.md_init()
    gm.md_step(5)
    x, _, _, _ = gm.state()
    assert _bond_residual(s, x) < 1e-4
    assert gm.thermo()["status_flags"] == 0
    gm.close()


@pytest.mark.timeout(600)
def test_salt_many_cells(ctx):
    """8,000 cells: binning without LDS aggregation and the atomic scatter"""
    _threads()
    s = make_salt(n_ions=64000, seed=9)
    r = salt_config(cut_off=2.5, r_core=2.5, alpha=1.44)
    print(f"[case] salt 64k, rc 2.5 n={s.n_atoms}", flush=True)
    g, o = GpuSystem(ctx, s, r), oracle.OracleSystem(s, r)
    assert int(np.prod(g.geometry()[0])) > 4096
    for a, b in zip(g.locator(), o.locator()):
        assert np.array_equal(a, b)
    gi, go, gn = g.neighbors(120)
    oi, oo, on = o.neighbors(120)
    assert np.array_equal(gn, on)
    assert np.array_equal(gi, oi)
    assert np.array_equal(go, oo)
    gf, gv, _, _ = g.force_real()
    of, ov, _, _ = o.force_real()
    assert rel_err(gf, of) < FP64_TOL
    assert rel_err(gv, ov) < FP64_TOL


def _alone(ctx, s, r, steps):
    g = GpuSystem(ctx, s, r)
    g.md_init()
    g.md_step(steps)
    out = g.state()
    g.close()
    return out


@pytest.mark.timeout(1200)
def test_ensemble_512_full_size(ctx):
    from rbmd_amd.lib import Context
    from rbmd_amd.serve import PackGroup, build_parallel, ensemble_specs

    _threads()
    specs = ensemble_specs(512, 500, seed=2024)
    print(f"[case] ensemble 512 systems, {sum(sp.atoms for sp in specs)} atoms", flush=True)
    inputs = build_parallel(specs, "mixed")
    ctxs = {k: Context(0) for k in ("salt", "water")}
    groups = {k: PackGroup(ctxs[k]) for k in ctxs}
    gs = []
    for i, (sp, (s, r)) in enumerate(zip(specs, inputs)):
        g = GpuSystem(ctxs[sp.kind], s, r)
        g.md_init()
        gs.append(g)
        groups[sp.kind].add(i, g, sp.steps)
    print("[case] ensemble built", flush=True)
    # every member stays in its pack through both rounds (finish() drops the
    # members that reach their count, so each round restarts the member lists)
    members = {k: [[i, gs[i], 0] for i, sp in enumerate(specs) if sp.kind == k] for k in ctxs}

    def step_all(n):
        for k, gr in groups.items():
            gr.members = [[i, g, n] for i, g, _ in members[k]]
            gr._drop_pack()
            while gr.members:
                gr.launch(50)
                gr.finish()

    step_all(10)
    # the smallest member of each kind against the oracle's FP64 run of 10 steps
    for kind in ("salt", "water"):
        k = min((i for i, sp in enumerate(specs) if sp.kind == kind), key=lambda i: specs[i].atoms)
        s, r = inputs[k]
        o = oracle.OracleSystem(s, type(r)(**{**r.__dict__, "precision": "fp64"}))
        assert o.md_init() == 0 and o.md_step(10) == 0
        xg, vg, fg, _ = gs[k].state()
        xo, vo, fo, _ = o.state()
        d = xg - xo
        d -= s.box * np.round(d / s.box)
        assert np.abs(d).max() < 1e-6 * s.box.max(), kind
        assert _rel_rms(vg, vo) < 1e-5 and _rel_rms(fg, fo) < 1e-5, kind
    step_all(490)
    for g in gs:
        assert g.thermo()["status_flags"] == 0
        assert g.thermo()["step"] == 500
    rng = np.random.default_rng(5)
    sample = sorted(set(rng.choice(len(specs), 6, replace=False).tolist()) | {0, len(specs) - 1})
    for k in sample:
        ref = _alone(ctx, *inputs[k], 500)
        for a, b in zip(ref, gs[k].state()):
            assert np.array_equal(a, b), k
    for gr in groups.values():
        gr.close()
    for g in gs:
        g.close()
    for c in ctxs.values():
        c.close()


@pytest.mark.timeout(1200)
def test_served_stream_rate_scaled(ctx):
    from rbmd_amd.serve import build, serve, stream_specs

    jobs = stream_specs(0.5, 20.0, seed=77)
    print(f"[case] served stream, {len(jobs)} jobs", flush=True)
    inputs = {j: build(sp, "mixed") for j, sp in enumerate(jobs)}
    results = {}
    r = serve(jobs, prebuilt=inputs, precision="mixed", results=results)
    assert r["jobs"] == len(jobs) == len(results) > 3
    assert r["atom_steps"] == sum(sp.atoms * sp.steps for sp in jobs)
    assert r["p99_latency_s"] >= r["p50_latency_s"] > 0
    for kind in ("salt", "water"):
        ks = [j for j, sp in enumerate(jobs) if sp.kind == kind]
        if not ks:
            continue
        j = min(ks, key=lambda i: jobs[i].atoms * jobs[i].steps)
        ref = _alone(ctx, *inputs[j], jobs[j].steps)
        for a, b in zip(ref, results[j]):
            assert np.array_equal(a, b), j
