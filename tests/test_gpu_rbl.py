"""Random-batch-list (RBL) path on the GPU against the oracle.

Water with the reference's own RBL settings (res/json_files/H2O/rbmd.json:
r_core 5, neighbor_sample_num 100, cut_off 10) and special bonds [0, 0, 0]:
- lists bit-exact (core region and every pice_num-th shell atom);
- the corrected near force ComputeNearForceRBLERFSpecialBondsWorklet
  (RunWorklet.cpp:1095-1266) with mean removal, in both erfc modes (the step
  table is the reference's ERFTable.h values), within 1e-10;
- the energies over the full rc list (ExecutionMD.cpp:670-750) within 1e-10;
- SpecialCoulGeneral (RunWorklet.cpp:2012-2084) in the bonded force;
- 20 MD steps within 1e-8 (state, thermo incl. energies and pressure), and a
  60-step run whose SHAKE moves put atoms off their binned cells, so the RBL
  misfit list kernel runs.
"""
import numpy as np
import pytest

import oracle
from rbmd_amd.lib import GpuSystem, RbmdError
from tests.systems import rel_err, salt, salt_run, water, water_run
from tests.test_gpu_parity import _md_compare, pair

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10
RBL_WATER = dict(neighbor_type="RBL", r_core=5.0, neighbor_sample_num=100)
RBL_SALT = dict(neighbor_type="RBL", r_core=2.0, neighbor_sample_num=40)


@pytest.mark.parametrize("mode", ["analytic", "table"])
def test_rbl_water_lists_force_energy(ctx, mode):
    s, r = water(), water_run(erfc_mode=mode, **RBL_WATER)
    g, o = pair(ctx, s, r)
    cap = 400
    gi, go, gn = g.neighbors(cap)
    oi, oo, on = o.neighbors(cap)
    assert np.array_equal(gn, on)
    assert np.array_equal(gi, oi)
    assert np.array_equal(go, oo)
    gf, gv, ge, gc = g.force_real()
    of, ov, oe, oc = o.force_real()
    assert rel_err(gf, of) < FP64_TOL
    assert rel_err(gv, ov) < 1e-9
    # energies: every atom, and the sums the thermo rows report
    assert np.abs(ge - oe).max() <= FP64_TOL * np.abs(oe).max()
    assert np.abs(gc - oc).max() <= FP64_TOL * np.abs(oc).max()
    assert abs(ge.sum() - oe.sum()) <= FP64_TOL * np.abs(oe).sum()
    assert abs(gc.sum() - oc.sum()) <= FP64_TOL * np.abs(oc).sum()
    assert np.abs(oc).sum() > 0 and np.abs(oe).sum() > 0


def test_rbl_water_special_coul(ctx):
    """bonded force + the intramolecular Coulomb correction; the correction is
    the bare Coulomb force between the atoms of each molecule, removed"""
    rng = np.random.default_rng(4)
    s = water()
    s = type(s)(**{**s.__dict__, "x": s.x + rng.uniform(-0.05, 0.05, s.x.shape)})
    r = water_run(**RBL_WATER)
    g, o = pair(ctx, s, r)
    gf, gb, ga = g.force_bonded()
    of, ob, oa = o.force_bonded()
    assert rel_err(gf, of) < FP64_TOL
    # the Verlet path has no such term: the difference is the correction alone
    gv = GpuSystem(ctx, s, water_run())
    fv, _, _ = gv.force_bonded()
    corr = gf - fv
    mol = s.mol_id
    x = s.x
    want = np.zeros_like(x)
    for i in range(s.n_atoms):
        for j in np.nonzero(mol == mol[i])[0]:
            if j == i:
                continue
            d = x[i] - x[j]
            d -= s.box * np.round(d / s.box)
            want[i] -= 332.06371 * s.charge[i] * s.charge[j] * d / np.linalg.norm(d) ** 3
    assert rel_err(corr, want) < 1e-12


@pytest.mark.parametrize("mode", ["analytic", "table"])
def test_rbl_water_md(ctx, mode):
    t = _md_compare(ctx, water(), water_run(erfc_mode=mode, **RBL_WATER), 20, 1e-8)
    assert t["status_flags"] == 0
    assert t["e_coul"] != 0.0 and t["e_lj"] != 0.0


def test_rbl_water_misfit_path(ctx):
    """over 60 steps SHAKE moves atoms off their binned cells and the RBL misfit
    kernel lists them from their own stencil; the trajectory still matches"""
    g, o = pair(ctx, water(), water_run(**RBL_WATER))
    g.md_init()
    assert o.md_init() == 0
    seen = 0
    for _ in range(6):
        g.md_step(10)
        assert o.md_step(10) == 0
        seen += g.thermo()["misfit_atoms"]
    xg, vg, fg, _ = g.state()
    xo, vo, fo, _ = o.state()
    assert rel_err(fg, fo) < 1e-7
    assert seen > 0
    tg, to = g.thermo(), o.thermo()
    for k in ("temperature", "e_lj", "e_coul"):
        assert abs(tg[k] - to[k]) <= 1e-7 * max(1.0, abs(to[k])), k


def test_rbl_salt_md(ctx):
    t = _md_compare(ctx, salt(n=2000, seed=6), salt_run(erfc_mode="table", **RBL_SALT), 20, 1e-8)
    assert t["status_flags"] == 0


def test_rbl_unsupported_combinations(ctx):
    from rbmd_amd.abi import RBMD_ERR_UNSUPPORTED

    s = salt(n=512, seed=1)
    with pytest.raises(RbmdError) as e:
        GpuSystem(ctx, s, salt_run(force_field="LJ/CUT", coulomb_type="NONE", **RBL_SALT))
    assert e.value.code == RBMD_ERR_UNSUPPORTED
    w = water(nmol=600, seed=5)
    with pytest.raises(RbmdError) as e:
        GpuSystem(ctx, w, water_run(cut=12.0, erfc_mode="table", **RBL_WATER))
    assert e.value.code == RBMD_ERR_UNSUPPORTED
