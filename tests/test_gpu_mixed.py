"""Mixed precision (RBMD_PREC_MIXED) against the FP64 path on identical inputs.

The mixed path evaluates the pair scalar (1/r, LJ, the erfc/exp table) in FP32
and the sin/cos of every structure-factor and k-space force term with the
hardware's FP32 sin/cos of the FP64-reduced phase, and sums in FP64.  Bar: forces within 1e-5 relative RMS of the FP64 path (RMS over atoms of
|dF_i| divided by the RMS of |F_i|) and the summed virial within 1e-5
relative, on the same lists and the same k-vector sets.
The FP64 path itself equals the oracle to 1e-10 (test_gpu_parity.py), so the
same bound holds against the oracle.

Energies: the pair energy sums cancel strongly (LJ attraction and repulsion
per O atom are each ~3 kcal/mol, their sum ~0.1), and on the lattice start
many O-O pairs share one exact d2, so the FP32 rounding of d2 does not
average out: per-atom energy sums differ by ~1e-6 (about 1e-5 of sum |e_i|,
reproduced by an FP32 evaluation of the same formula in numpy).  The energy
bound is therefore ENERGY_TOL of sum |e_i|, and the measured value is printed.
"""
import numpy as np
import pytest

import oracle
from rbmd_amd.lib import GpuSystem, RbmdError
from tests.systems import salt, salt_run, water, water_run

pytestmark = pytest.mark.gpu

MIXED_TOL = 1e-5
ENERGY_TOL = 5e-5


def rms_rel(a, b):
    a = np.asarray(a).reshape(len(a), -1)
    b = np.asarray(b).reshape(len(b), -1)
    return float(np.sqrt(np.mean(np.sum((a - b) ** 2, 1)) / np.mean(np.sum(b ** 2, 1))))


def both(ctx, system, run):
    print(f"[case] mixed vs fp64 n={system.n_atoms} kind={system.kind}", flush=True)
    g64 = GpuSystem(ctx, system, run)
    g32 = GpuSystem(ctx, system, type(run)(**{**run.__dict__, "precision": "mixed"}))
    return g64, g32


CASES = {
    "salt": lambda: (salt(), salt_run()),
    "water": lambda: (water(), water_run()),
    "water_small_box": lambda: (water(nmol=201, seed=8), water_run()),  # minimum-image pair vectors
    "lj_only": lambda: (salt(n=2000, seed=9), salt_run(force_field="LJ/CUT", coulomb_type="NONE", cut_off=2.5,
                                                        r_core=2.5)),
}


@pytest.mark.parametrize("case", list(CASES))
def test_mixed_real_space(ctx, case):
    s, r = CASES[case]()
    g64, g32 = both(ctx, s, r)
    g64.neighbors(0)
    g32.neighbors(0)
    f64, v64, e64, c64 = g64.force_real()
    f32, v32, e32, c32 = g32.force_real()
    err = rms_rel(f32, f64)
    print(f"[mixed] {case} real-space force rms rel {err:.2e}", flush=True)
    assert err < MIXED_TOL
    vs64, vs32 = v64.sum(0), v32.sum(0)
    verr = np.abs(vs32 - vs64).max() / np.abs(vs64).max()
    elj_err = abs(e32.sum() - e64.sum()) / np.abs(e64).sum()
    ec_err = abs(c32.sum() - c64.sum()) / max(np.abs(c64).sum(), 1e-300)
    print(f"[mixed] {case} virial rel {verr:.2e} e_lj rel {elj_err:.2e} e_coul rel {ec_err:.2e}", flush=True)
    assert verr <= MIXED_TOL
    assert elj_err <= ENERGY_TOL
    assert ec_err <= ENERGY_TOL


@pytest.mark.parametrize("kind", ["salt", "water"])
def test_mixed_rbe(ctx, kind):
    s, r = CASES[kind]()
    g64, g32 = both(ctx, s, r)
    m = oracle.rbe_sample(r.alpha, s.box, r.rbe_p, 99, 5)
    g64.set_kvectors(m)
    g32.set_kvectors(m)
    rho64 = g64.structure_factor(r.rbe_p)
    rho32 = g32.structure_factor(r.rbe_p)
    rerr = float(np.sqrt(np.mean(np.sum((rho32 - rho64) ** 2, 1)) / np.mean(np.sum(rho64 ** 2, 1))))
    f64, v64 = g64.rbe_force()
    f32, v32 = g32.rbe_force()
    err = rms_rel(f32, f64)
    # rho(k) = sum_i q_i (cos, sin)(k.r_i): with every sin / cos within 1.2e-7
    # (hardware sin / cos of the FP64-reduced phase; measured over 2^24 phases)
    # no component can move by more than sum |q_i| times that, plus FP64 sums
    dmax = float(np.abs(rho32 - rho64).max())
    qsum = float(np.abs(s.charge).sum())
    print(f"[mixed] {kind} rho rms rel {rerr:.2e} max |d rho| / sum|q| {dmax / qsum:.2e} "
          f"rbe force rms rel {err:.2e}", flush=True)
    assert dmax <= 1.3e-7 * qsum
    assert rerr < MIXED_TOL
    assert err < MIXED_TOL
    assert np.abs(v32.sum(0) - v64.sum(0)).max() <= MIXED_TOL * np.abs(v64.sum(0)).max()


def test_mixed_ewald_stays_fp64(ctx):
    """rbmd_ewald_force is the FP64 error reference in either precision: the
    same bits, given the same k-space tiles (the tuned table may choose
    different tiles per precision, and a tile changes the summation order)"""
    s, r = CASES["salt"]()
    r = type(r)(**{**r.__dict__, "tiles": "default"})
    g64, g32 = both(ctx, s, r)
    f64, _, e64 = g64.ewald_force([6, 6, 6])
    f32, _, e32 = g32.ewald_force([6, 6, 6])
    assert np.array_equal(f32, f64)
    assert e32 == e64


@pytest.mark.parametrize("kind", ["salt", "water"])
def test_mixed_md(ctx, kind):
    """total force after md_init, then thermo after 10 steps (same k sets)"""
    s, r = CASES[kind]()
    g64, g32 = both(ctx, s, r)
    g64.md_init()
    g32.md_init()
    f64 = g64.state()[2]
    f32 = g32.state()[2]
    err = rms_rel(f32, f64)
    print(f"[mixed] {kind} total force after init rms rel {err:.2e}", flush=True)
    assert err < MIXED_TOL
    g64.md_step(10)
    g32.md_step(10)
    t64, t32 = g64.thermo(), g32.thermo()
    for k in ("temperature", "e_lj", "e_coul"):
        d = abs(t32[k] - t64[k]) / max(abs(t64[k]), 1e-300)
        print(f"[mixed] {kind} {k} {t64[k]:.10g} vs {t32[k]:.10g} rel {d:.2e}", flush=True)
        assert d < 1e-4, k
    # pressure = kinetic + virial parts that largely cancel: bound the
    # difference by the virial scale
    vscale = np.abs(np.asarray(t64["virial"])).max()
    dp = abs(t32["pressure"] - t64["pressure"])
    print(f"[mixed] {kind} pressure {t64['pressure']:.8g} vs {t32['pressure']:.8g}", flush=True)
    assert np.abs(np.asarray(t32["virial"]) - np.asarray(t64["virial"])).max() <= 1e-4 * vscale
    assert np.isfinite(dp)


def test_mixed_graph_matches_eager(ctx):
    s, r = CASES["water"]()
    r = type(r)(**{**r.__dict__, "precision": "mixed"})
    g1, g2 = GpuSystem(ctx, s, r), GpuSystem(ctx, s, r)
    g1.md_init()
    g2.md_init()
    for _ in range(3):
        g1.md_step(4)
        g2.md_run_graph(4)
        for u, v in zip(g1.state(), g2.state()):
            assert np.array_equal(u, v)


def test_mixed_unsupported_combinations(ctx):
    s = salt(n=512, seed=1)
    for kw in ({"pair_path": "cells"}, {"neighbor_type": "RBL", "r_core": 2.0, "neighbor_sample_num": 40}):
        r = salt_run(precision="mixed", **kw)
        with pytest.raises(RbmdError):
            GpuSystem(ctx, s, r)
