"""Pins of the CPU FP64 oracle against what the mathematics fixes (CPU only).

The oracle restates the reference algorithms; nothing in the reference can be
run here, so each part is checked against a closed form, an invariant, a
brute-force enumeration or a textbook constant, chosen so that a plausible
slip (a dropped term, a wrong sign, a transposed index) fails one of them:

  LJ pair                closed form for two atoms (force and energy)
  erfc-Coulomb pair      closed form G(r) with math.erfc for two charges
  real-space total       finite differences of the energy against the force
  neighbour lists        all-image brute-force counts; every entry valid, no duplicates
  RBE S                  theta-function identity S = H^3 - 1 vs the direct k sum,
                         and the salt box's value 734.5889
  RBE force              momentum conservation; unbiased against the Ewald
                         k-space force (average over independent k sets)
  k sampler              marginal frequencies against the exact discrete Gaussian
  Ewald / Madelung       rock-salt Madelung constant 1.747564594633
  integrator             time reversibility of velocity Verlet
  SHAKE / RATTLE         bond-length and bond-velocity residuals
"""
import math

import numpy as np
import pytest

import oracle
from rbmd_amd import RunConfig, make_salt, make_water, rocksalt_lattice, salt_config, water_config
from rbmd_amd.system import System


def two_atoms(r, q=(0.0, 0.0), eps=1.0, sigma=1.0, L=20.0):
    x = np.array([[5.0, 7.0, 9.0], [5.0 + r, 7.0, 9.0]])
    return System(lo=np.zeros(3), hi=np.full(3, L), x=x, v=np.zeros((2, 3)), type=np.zeros(2, np.int32),
                  charge=np.array(q, float), mass=np.ones(2), eps=np.array([eps]), sigma=np.array([sigma]))


def test_lj_two_atoms_closed_form():
    r = 1.3
    o = oracle.OracleSystem(two_atoms(r), salt_config(force_field="LJ/CUT", coulomb_type="NONE", cut_off=4.0))
    o.neighbors(0)
    f, _, elj, _ = o.force_real()
    s6 = (1.0 / r) ** 6
    # F_0 = 24 (2 s^12 - s^6) / r^2 * (x_0 - x_1)
    fx0 = 24.0 * (2 * s6 * s6 - s6) / r ** 2 * (-r)
    assert abs(f[0, 0] - fx0) <= 1e-13 * abs(fx0)
    assert abs(f[1, 0] + fx0) <= 1e-13 * abs(fx0)
    assert np.abs(f[:, 1:]).max() == 0.0
    u = 4.0 * (s6 * s6 - s6)
    assert abs(elj.sum() - u) <= 1e-13 * abs(u)


def test_erfc_coulomb_two_charges_closed_form():
    r, beta = 1.7, 0.8
    run = salt_config(alpha=beta * beta, cut_off=4.0)
    o = oracle.OracleSystem(two_atoms(r, q=(1.0, -1.0), eps=0.0), run)
    o.neighbors(0)
    f, _, _, ec = o.force_real()
    G = math.erfc(beta * r) / r ** 2 + 2 * beta / math.sqrt(math.pi) * math.exp(-(beta * r) ** 2) / r
    qq = -1.0
    fx0 = -qq * G  # attraction: atom 0 is pulled towards +x
    assert abs(f[0, 0] - fx0) <= 1e-13 * abs(fx0)
    assert abs(f[1, 0] + fx0) <= 1e-13 * abs(fx0)
    e = qq * math.erfc(beta * r) / r
    assert abs(ec.sum() - e) <= 1e-13 * abs(e)


def _energy(o, x):
    o.set_state(x=x)
    o.neighbors(0)
    _, _, elj, ec = o.force_real()
    return elj.sum() + ec.sum()


def test_real_space_force_is_minus_energy_gradient():
    s = make_salt(n_ions=500, seed=21)
    o = oracle.OracleSystem(s, salt_config())
    o.neighbors(0)
    f, _, _, _ = o.force_real()
    h = 1e-6
    rng = np.random.default_rng(0)
    for a in rng.choice(s.n_atoms, 6, replace=False):
        for d in range(3):
            xp = s.x.copy()
            xp[a, d] += h
            xm = s.x.copy()
            xm[a, d] -= h
            g = (_energy(o, xp) - _energy(o, xm)) / (2 * h)
            assert abs(-g - f[a, d]) <= 2e-6 * max(1.0, abs(f[a, d])), (a, d, -g, f[a, d])


@pytest.mark.parametrize("kind", ["salt", "water"])
def test_neighbor_lists_against_bruteforce(kind):
    if kind == "salt":
        s, r = make_salt(n_ions=1000, seed=7), salt_config()
    else:
        s, r = make_water(n_mol=300, seed=4), water_config()
    o = oracle.OracleSystem(s, r)
    ids, off, num = o.neighbors(0)
    assert np.array_equal(num, o.bruteforce_counts(r.cut_off))
    rc2 = r.cut_off ** 2
    for i in range(0, s.n_atoms, 7):
        k = num[i]
        d = (s.x[ids[i, :k]] - off[i, :k]) - s.x[i]  # offset = dxdydz (wrapped - unwrapped)
        d2 = (d * d).sum(1)
        assert np.all(rc2 - d2 > 1e-4) and np.all(d2 > 1e-4)
        keys = set(zip(ids[i, :k].tolist(), map(tuple, np.round(off[i, :k], 6).tolist())))
        assert len(keys) == k  # each (neighbour, image) once


@pytest.mark.parametrize("alpha,L", [(0.64, 20.0), (0.09, 66.88), (1.69, 8.0)])
def test_rbe_S_theta_identity(alpha, L):
    box = np.full(3, L)
    S = oracle.rbe_S(alpha, box)
    assert abs(S - oracle.rbe_S_direct(alpha, box, 60)) <= 1e-12 * S


def test_rbe_S_salt_box_value():
    assert abs(oracle.rbe_S(0.64, np.full(3, 20.0)) - 734.5888522) < 1e-6


def test_rbe_force_conserves_momentum():
    s = make_salt(n_ions=1000, seed=2)
    r = salt_config()
    o = oracle.OracleSystem(s, r)
    for stream in (1, 2, 3):
        o.set_kvectors(oracle.rbe_sample(r.alpha, s.box, r.rbe_p, 99, stream))
        f, _ = o.rbe_force()
        assert np.abs(f.sum(0)).max() <= 1e-11 * np.abs(f).sum()


def test_rbe_force_unbiased_against_ewald():
    """E[F_RBE] = F_Ewald,k: the average over M independent k sets converges
    like 1/sqrt(M) to the k-space Ewald force at the same alpha."""
    s = make_salt(n_ions=216, seed=5, density=0.5)
    r = salt_config(rbe_p=100)
    o = oracle.OracleSystem(s, r)
    fe, _, _ = o.ewald_force([14, 14, 14])
    acc = np.zeros_like(fe)
    errs = {}
    for m in range(1, 401):
        o.set_kvectors(oracle.rbe_sample(r.alpha, s.box, r.rbe_p, 2024, m))
        f, _ = o.rbe_force()
        acc += f
        if m in (25, 400):
            errs[m] = np.sqrt(((acc / m - fe) ** 2).sum(1).mean()) / np.sqrt((fe ** 2).sum(1).mean())
    assert errs[400] < 0.5 * errs[25]
    assert errs[400] < 0.3


def test_sampler_marginal_distribution():
    alpha, L = 0.64, 20.0
    m = oracle.rbe_sample(alpha, np.full(3, L), 20000, 7, 1)
    assert np.all(np.abs(m).sum(1) > 0)  # the zero vector is never drawn
    g = lambda k: np.exp(-(math.pi ** 2) * k * k / (alpha * L * L))
    ks = np.arange(-40, 41)
    H = g(ks).sum()
    for a in range(3):
        col = m[:, a].astype(int)
        for k in range(0, 5):
            p = (g(k) * H * H - (1.0 if k == 0 else 0.0)) / (H ** 3 - 1)
            p = p if k == 0 else 2 * p
            obs = np.mean(np.abs(col) == k)
            sd = math.sqrt(p * (1 - p) / len(col))
            assert abs(obs - p) < 12 * sd + 1e-3, (a, k, obs, p)


def test_madelung_constant():
    lat = rocksalt_lattice(8)
    beta = 1.3
    run = RunConfig(unit="LJ", force_field="LJ/CUT/COUL/LONG", cut_off=3.99, r_core=3.99, coulomb_type="EWALD",
                    alpha=beta * beta, kmax=[22, 22, 22], temperature=[1, 1, 1])
    o = oracle.OracleSystem(lat, run)
    o.neighbors(0)
    _, _, _, ec = o.force_real()
    _, _, ek = o.ewald_force([22, 22, 22])
    self_e = -math.sqrt(run.alpha / math.pi) * np.sum(lat.charge ** 2)
    E = ec.sum() + ek + self_e
    assert abs(-2 * E / lat.n_atoms - 1.747564594633) < 1e-10


def test_velocity_verlet_time_reversible():
    s = make_salt(n_ions=500, seed=8)
    run = salt_config(force_field="LJ/CUT", coulomb_type="NONE", cut_off=2.5, r_core=2.5, temp_ctrl="NONE")
    o = oracle.OracleSystem(s, run)
    assert o.md_init() == 0
    x0, v0, _, _ = o.state()
    assert o.md_step(40) == 0
    x1, v1, _, _ = o.state()
    assert np.abs(x1 - x0).max() > 1e-3
    o2 = oracle.OracleSystem(s, run)
    o2.set_state(x=x1, v=-v1)
    assert o2.md_init() == 0
    assert o2.md_step(40) == 0
    x2, v2, _, _ = o2.state()
    d = x2 - x0
    d -= s.box * np.round(d / s.box)
    assert np.abs(d).max() < 1e-9
    assert np.abs(v2 + v0).max() < 1e-9


def test_shake_rattle_constraints():
    s = make_water(n_mol=300, seed=6)
    o = oracle.OracleSystem(s, water_config())
    assert o.md_init() == 0
    assert o.md_step(20) == 0
    x, v, _, _ = o.state()
    ang = s.angle_atoms.reshape(-1, 3)  # H, O, H
    box = s.box

    def vec(a, b):
        d = x[b] - x[a]
        return d - box * np.round(d / box)

    r1 = np.linalg.norm(vec(ang[:, 1], ang[:, 0]), axis=1)
    r2 = np.linalg.norm(vec(ang[:, 1], ang[:, 2]), axis=1)
    rhh = np.linalg.norm(vec(ang[:, 0], ang[:, 2]), axis=1)
    assert np.abs(r1 - 1.0).max() < 1e-4 and np.abs(r2 - 1.0).max() < 1e-4
    assert np.abs(rhh - 2 * math.sin(math.radians(109.47 / 2))).max() < 1e-4
    # RATTLE: no relative velocity along the constrained bonds
    d = vec(ang[:, 1], ang[:, 0])
    dv = v[ang[:, 0]] - v[ang[:, 1]]
    assert np.abs((d * dv).sum(1)).max() < 1e-6 * np.abs(v).max()


def test_bonded_forces_are_minus_energy_gradient():
    """harmonic bonds E = k (r - r0)^2 and angles E = k (theta - theta0)^2
    (MolecularWorklet.h:58-397): forces against finite differences of the
    bonded energy on a jittered water box"""
    s = make_water(n_mol=50, seed=9)
    rng = np.random.default_rng(3)
    x0 = s.x + rng.uniform(-0.08, 0.08, s.x.shape)
    o = oracle.OracleSystem(s, water_config())
    o.set_state(x=x0)
    f, eb, ea = o.force_bonded()
    assert eb > 0 and ea > 0

    def energy(x):
        o.set_state(x=x)
        _, b, a = o.force_bonded()
        return b + a

    h = 1e-6
    for a in (0, 1, 2, 31, 77):
        for d in range(3):
            xp = x0.copy()
            xp[a, d] += h
            xm = x0.copy()
            xm[a, d] -= h
            g = (energy(xp) - energy(xm)) / (2 * h)
            assert abs(-g - f[a, d]) <= 1e-6 * max(1.0, abs(f[a, d])), (a, d, -g, f[a, d])
    # a single bond at r: the bond energy is k (r - r0)^2 exactly
    i, j = s.bond_atoms[0], s.bond_atoms[1]
    r = np.linalg.norm(x0[i] - x0[j])
    o2 = oracle.OracleSystem(s, water_config())
    x1 = s.x.copy()
    x1[j] = x1[i] + (x0[j] - x0[i])  # only bond 0 stretched
    o2.set_state(x=x1)
    _, eb1, _ = o2.force_bonded()
    assert abs(eb1 - s.bond_k[0] * (r - s.bond_r0[0]) ** 2) <= 1e-12 * max(eb1, 1e-30)
