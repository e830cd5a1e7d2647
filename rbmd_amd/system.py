"""Atom systems: the data that `init_configuration` produces in the reference.

* ``System`` holds what ModelFileInitCondition puts into Para
  (framework/InitCondition/ModelFileInitCondition.cpp:406-478, 736-956).
* ``read_data`` parses LAMMPS data files with atom_style atomic / charge /
  full, Masses, Pair Coeffs, Bond/Angle Coeffs, Bonds, Angles and Velocities
  (ModelFileInitCondition.cpp:120-278, 406-655).  Image flags on Atoms lines
  are ignored, as in the reference (:445-446).
* ``special_lists`` builds the 1-2/1-3/1-4 special-bond CSR lists exactly as
  SetSpecialBonds does (ModelFileInitCondition.cpp:848-956).
* ``make_salt`` / ``make_water`` generate the seeded synthetic systems of the
  charged-LJ salt and rigid 3-site water workloads.
"""
from __future__ import annotations

import dataclasses
import math
from typing import Optional

import numpy as np

WATER_MASS_O = 15.9994
WATER_MASS_H = 1.008
AVOGADRO = 6.02214076e23


@dataclasses.dataclass
class System:
    lo: np.ndarray
    hi: np.ndarray
    x: np.ndarray  # [n,3]
    v: np.ndarray  # [n,3]
    type: np.ndarray  # [n] 0-based
    charge: np.ndarray
    mass: np.ndarray
    eps: np.ndarray  # per type
    sigma: np.ndarray  # per type
    mol_id: Optional[np.ndarray] = None
    special_offsets: Optional[np.ndarray] = None
    special_ids: Optional[np.ndarray] = None
    special_weights: Optional[np.ndarray] = None
    bond_atoms: Optional[np.ndarray] = None
    bond_k: Optional[np.ndarray] = None
    bond_r0: Optional[np.ndarray] = None
    angle_atoms: Optional[np.ndarray] = None
    angle_k: Optional[np.ndarray] = None
    angle_theta0: Optional[np.ndarray] = None
    kind: str = "generic"

    @property
    def n_atoms(self) -> int:
        return int(self.x.shape[0])

    @property
    def box(self) -> np.ndarray:
        return np.asarray(self.hi, dtype=np.float64) - np.asarray(self.lo, dtype=np.float64)

    @property
    def volume(self) -> float:
        b = self.box
        return float(b[0] * b[1] * b[2])


def special_lists(n_atoms: int, bond_atoms: np.ndarray, special_bonds) -> tuple:
    """SetSpecialBonds (ModelFileInitCondition.cpp:848-956).

    The bond multimap keeps insertion order per key (std::unordered_multimap
    equal_range order is unspecified; insertion order is used here and the
    weight lookup is order independent unless an id repeats).
    Weights are ints in the reference (PARA_SPECIAL_BONDS is vector<int>).
    """
    w = [int(special_bonds[0]), int(special_bonds[1]), int(special_bonds[2])]
    links = [[] for _ in range(n_atoms)]
    nb = len(bond_atoms) // 2
    for b in range(nb):
        i, j = int(bond_atoms[2 * b]), int(bond_atoms[2 * b + 1])
        links[i].append(j)
        links[j].append(i)
    offsets = [0]
    ids = []
    weights = []
    for a in range(n_atoms):
        if not links[a]:
            ids.append(a)
            weights.append(1.0)
            offsets.append(offsets[-1] + 1)
            continue
        cnt = 0
        for k1 in links[a]:
            ids.append(k1)
            weights.append(float(w[0]))
            cnt += 1
            for k2 in links[k1]:
                if k2 == a:
                    continue
                ids.append(k2)
                weights.append(float(w[1]))
                cnt += 1
                for k3 in links[k2]:
                    if k3 == k1:
                        continue
                    ids.append(k3)
                    weights.append(float(w[2]))
                    cnt += 1
        offsets.append(offsets[-1] + cnt)
    return (np.asarray(offsets, np.int32), np.asarray(ids, np.int32), np.asarray(weights, np.float64))


def maxwell(rng: np.random.Generator, mass: np.ndarray, temperature: float, kB: float, dof: int,
            mvv2e: float = 1.0) -> np.ndarray:
    """Maxwell velocities, COM momentum removed, rescaled to exactly `temperature`."""
    n = len(mass)
    sig = np.sqrt(kB * temperature / (mass * mvv2e))
    v = rng.standard_normal((n, 3)) * sig[:, None]
    p = (mass[:, None] * v).sum(0) / mass.sum()
    v -= p
    ke2 = float((mass[:, None] * v * v).sum()) * mvv2e
    t_now = ke2 / (dof * kB)
    v *= math.sqrt(temperature / t_now)
    return v


def make_salt(n_ions: int = 4000, density: float = 0.5, seed: int = 1, temperature: float = 1.5,
              jitter: float = 0.05) -> System:
    """Charged-LJ salt on a rock-salt simple-cubic lattice (LJ units).

    Sites: m^3 with m = ceil(cbrt(n)) (even), charge (-1)^(i+j+k); the surplus
    sites are removed in equal numbers from each parity so the system stays
    neutral; uniform jitter of +-jitter per coordinate; masses 1; Maxwell
    velocities at `temperature` with dof 3N-3.  With n=4000 and density 0.5
    this is 16^3 sites in L = 20 (spacing 1.25), 48 sites of each parity removed.
    """
    if n_ions % 2:
        raise ValueError("n_ions must be even")
    rng = np.random.default_rng(seed)
    L = round((n_ions / density) ** (1.0 / 3.0), 10)  # exact box edge when n/density is a cube
    m = int(math.ceil(round(n_ions ** (1.0 / 3.0), 9)))
    if m % 2:
        m += 1
    a = L / m
    idx = np.arange(m)
    I, J, K = np.meshgrid(idx, idx, idx, indexing="ij")
    I, J, K = I.ravel(), J.ravel(), K.ravel()
    par = (I + J + K) % 2
    even = np.flatnonzero(par == 0)
    odd = np.flatnonzero(par == 1)
    drop = (m ** 3 - n_ions) // 2
    keep_even = np.sort(rng.choice(even, size=len(even) - drop, replace=False))
    keep_odd = np.sort(rng.choice(odd, size=len(odd) - drop, replace=False))
    keep = np.sort(np.concatenate([keep_even, keep_odd]))
    x = (np.stack([I[keep], J[keep], K[keep]], 1).astype(np.float64) + 0.5) * a
    x += rng.uniform(-jitter, jitter, size=x.shape)
    x = np.mod(x, L)
    q = np.where(par[keep] == 0, 1.0, -1.0)
    typ = np.where(q > 0, 0, 1).astype(np.int32)
    n = len(keep)
    mass = np.ones(n)
    v = maxwell(rng, mass, temperature, 1.0, 3 * n - 3)
    return System(lo=np.zeros(3), hi=np.full(3, L), x=x, v=v, type=typ, charge=q, mass=mass,
                  eps=np.array([1.0, 1.0]), sigma=np.array([1.0, 1.0]), kind="salt")


def make_inbuild(dims, x_range, y_range, z_range, velocity_type: str = "GAUSS", seed: int = 1) -> System:
    """The reference's in-build LJ lattice (MeshFreeCondition + LJInitCondition,
    framework/InitCondition/MeshFreeCondition.cpp:45-100, LJInitCondition.cpp:37-99):
    one particle at the centroid of every cell of a uniform dims[0] x dims[1] x
    dims[2] grid over the ranges (x fastest, as VTK-m orders structured cells),
    mass 1, one type with eps = sigma = 1, no charge; each atom's special list
    is itself with weight 1.  GAUSS velocities are standard normal per
    component (the reference draws them from an unseeded generator; here from
    `seed`); ZERO gives zero velocities."""
    dims = [int(d) for d in dims]
    lo = np.array([x_range[0], y_range[0], z_range[0]], np.float64)
    hi = np.array([x_range[1], y_range[1], z_range[1]], np.float64)
    space = (hi - lo) / np.asarray(dims, np.float64)
    k, j, i = np.meshgrid(np.arange(dims[2]), np.arange(dims[1]), np.arange(dims[0]), indexing="ij")
    cells = np.stack([i.ravel(), j.ravel(), k.ravel()], 1).astype(np.float64)
    x = lo + (cells + 0.5) * space
    n = len(x)
    if velocity_type == "GAUSS":
        v = np.random.default_rng(seed).standard_normal((n, 3))
    else:
        v = np.zeros((n, 3))
    ids = np.arange(n, dtype=np.int32)
    return System(lo=lo, hi=hi, x=x, v=v, type=np.zeros(n, np.int32), charge=np.zeros(n), mass=np.ones(n),
                  eps=np.array([1.0]), sigma=np.array([1.0]), mol_id=ids.copy(),
                  special_offsets=np.arange(n + 1, dtype=np.int32), special_ids=ids,
                  special_weights=np.ones(n), kind="inbuild")


def water_geometry(rng: np.random.Generator, n: int, bond: float = 1.0, angle_deg: float = 109.47):
    """Random orientations of an H-O-H unit: returns H1, H2 offsets from O, [n,3] each."""
    half = math.radians(angle_deg) / 2.0
    # body frame: O at origin, H in the xz plane
    h1 = np.array([math.sin(half), 0.0, math.cos(half)]) * bond
    h2 = np.array([-math.sin(half), 0.0, math.cos(half)]) * bond
    qv = rng.standard_normal((n, 4))
    qv /= np.linalg.norm(qv, axis=1, keepdims=True)
    w, x, y, z = qv.T
    R = np.empty((n, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z)
    R[:, 0, 1] = 2 * (x * y - z * w)
    R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w)
    R[:, 1, 1] = 1 - 2 * (x * x + z * z)
    R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w)
    R[:, 2, 1] = 2 * (y * z + x * w)
    R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R @ h1, R @ h2


def rigidify_velocities(x: np.ndarray, v: np.ndarray, mass: np.ndarray, L: float, n_mol: int) -> np.ndarray:
    """Project every molecule's velocities onto rigid-body motion (COM + rotation),
    all molecules at once (atoms 3k, 3k+1, 3k+2 form molecule k)."""
    r = x.reshape(n_mol, 3, 3).copy()
    r[:, 1:] -= L * np.round((r[:, 1:] - r[:, :1]) / L)
    vv = v.reshape(n_mol, 3, 3)
    m = mass.reshape(n_mol, 3)
    M = m.sum(1)
    rc = (m[:, :, None] * r).sum(1) / M[:, None]
    vc = (m[:, :, None] * vv).sum(1) / M[:, None]
    d = r - rc[:, None, :]
    Lang = (m[:, :, None] * np.cross(d, vv - vc[:, None, :])).sum(1)
    dd = (d * d).sum(2)
    I = (m * dd).sum(1)[:, None, None] * np.eye(3) - np.einsum("mk,mki,mkj->mij", m, d, d)
    om = np.linalg.solve(I, Lang[:, :, None])[:, :, 0]
    return (vc[:, None, :] + np.cross(om[:, None, :], d)).reshape(-1, 3)


def water_special_lists(n_mol: int, special_bonds) -> tuple:
    """special_lists() for molecules bonded O-H1, O-H2 (bonds listed in that
    order), built directly: O lists H1, H2 (1-2); each H lists O (1-2) and the
    other H (1-3).  Same order and weights as the generic walk."""
    w0, w1 = float(int(special_bonds[0])), float(int(special_bonds[1]))
    base = 3 * np.arange(n_mol)
    ids = np.stack([base + 1, base + 2, base, base + 2, base, base + 1], 1).reshape(-1).astype(np.int32)
    w = np.tile(np.array([w0, w0, w0, w1, w0, w1]), n_mol)
    off = (2 * np.arange(3 * n_mol + 1)).astype(np.int32)
    return off, ids, w


def make_water(n_mol: int = 10000, density_g_cm3: float = 1.0, seed: int = 1, temperature: float = 300.0,
               q_o: float = -0.8476, min_dist: float = 1.7) -> System:
    """Rigid 3-site water in REAL units.

    Molecules sit on the first n_mol sites (seeded choice) of an m^3 cubic
    lattice filling the box at the given density, with random orientations
    drawn until no intermolecular atom pair is closer than `min_dist` A.
    Atom order per molecule: O, H, H (ids 3k, 3k+1, 3k+2); angle triplets are
    (H, O, H) as SHAKE expects (MolecularWorklet.h:2398-2400).
    Bonds k=540.634 r0=1.0, angle k=50 theta0=109.47 (examples/H2O/H2O_212121_v.data:14-30).
    """
    rng = np.random.default_rng(seed)
    mol_mass = WATER_MASS_O + 2 * WATER_MASS_H
    vol_cm3 = n_mol * mol_mass / AVOGADRO / density_g_cm3
    L = (vol_cm3 * 1e24) ** (1.0 / 3.0)
    m = int(math.ceil(n_mol ** (1.0 / 3.0) - 1e-9))
    a = L / m
    idx = np.arange(m)
    I, J, K = np.meshgrid(idx, idx, idx, indexing="ij")
    sites = np.stack([I.ravel(), J.ravel(), K.ravel()], 1)
    chosen = np.sort(rng.choice(len(sites), size=n_mol, replace=False))
    O = (sites[chosen].astype(np.float64) + 0.5) * a
    h1, h2 = water_geometry(rng, n_mol)
    # reject orientations with close intermolecular contacts (cell grid of the lattice)
    grid = -np.ones((m, m, m), dtype=np.int64)
    sc = sites[chosen]
    grid[sc[:, 0], sc[:, 1], sc[:, 2]] = np.arange(n_mol)
    offsets = [(dx, dy, dz) for dx in (-1, 0, 1) for dy in (-1, 0, 1) for dz in (-1, 0, 1) if (dx, dy, dz) != (0, 0, 0)]

    def contact(ia, ib):
        """closest atom distance between molecules ia[k] and ib[k] (minimum image);
        exactly symmetric in (ia, ib)"""
        pa = np.stack([O[ia], O[ia] + h1[ia], O[ia] + h2[ia]], 1)
        pb = np.stack([O[ib], O[ib] + h1[ib], O[ib] + h2[ib]], 1)
        d = pa[:, :, None, :] - pb[:, None, :, :]
        d -= L * np.round(d / L)
        return np.sqrt((d * d).sum(-1)).reshape(len(ia), -1).min(1)

    def flag_pairs(idx, offs):
        """molecules in a contact below min_dist among the pairs (i, lattice
        neighbour of i at offset o) for i in idx, o in offs; flags both sides"""
        bad = np.zeros(n_mol, bool)
        for dx, dy, dz in offs:
            nb = grid[(sc[idx, 0] + dx) % m, (sc[idx, 1] + dy) % m, (sc[idx, 2] + dz) % m]
            ii = np.flatnonzero(nb >= 0)
            if len(ii) == 0:
                continue
            hit = contact(idx[ii], nb[ii]) < min_dist
            bad[idx[ii][hit]] = True
            bad[nb[ii][hit]] = True
        return np.flatnonzero(bad)

    # every lattice-neighbour pair once (half of the offsets); then, after the
    # flagged molecules are re-oriented, only pairs touching them can change
    # (the test is symmetric and all other pairs were already clear)
    bad = flag_pairs(np.arange(n_mol), offsets[len(offsets) // 2:])
    for _ in range(199):
        if len(bad) == 0:
            break
        nh1, nh2 = water_geometry(rng, len(bad))
        h1[bad] = nh1
        h2[bad] = nh2
        bad = flag_pairs(bad, offsets)
    x = np.stack([O, O + h1, O + h2], 1).reshape(-1, 3)
    x = np.mod(x, L)
    n = 3 * n_mol
    typ = np.tile(np.array([0, 1, 1], np.int32), n_mol)
    q = np.tile(np.array([q_o, -q_o / 2, -q_o / 2]), n_mol)
    mass = np.tile(np.array([WATER_MASS_O, WATER_MASS_H, WATER_MASS_H]), n_mol)
    mol = np.repeat(np.arange(n_mol, dtype=np.int32), 3)
    base = 3 * np.arange(n_mol)
    bonds = np.stack([base, base + 1, base, base + 2], 1).reshape(-1).astype(np.int32)
    angles = np.stack([base + 1, base, base + 2], 1).reshape(-1).astype(np.int32)
    kB = 1.9872067e-3
    mvv2e = 1.0 / 4.186e-4
    dof = 3 * n - n - 3  # ComputeTempe with fix_shake: 3n - n - 3 (ExecutionNVT.cpp:764-767)
    v = rng.standard_normal((n, 3)) * np.sqrt(kB * temperature / (mass * mvv2e))[:, None]
    v = rigidify_velocities(x, v, mass, L, n_mol)
    v -= (mass[:, None] * v).sum(0) / mass.sum()
    t_now = float((mass[:, None] * v * v).sum()) * mvv2e / (dof * kB)
    v *= math.sqrt(temperature / t_now)
    so, si, sw = water_special_lists(n_mol, [0, 0, 0])
    return System(lo=np.zeros(3), hi=np.full(3, L), x=x, v=v, type=typ, charge=q, mass=mass,
                  eps=np.array([0.1553, 0.0]), sigma=np.array([3.166, 0.0]), mol_id=mol,
                  special_offsets=so, special_ids=si, special_weights=sw,
                  bond_atoms=bonds, bond_k=np.full(2 * n_mol, 540.634), bond_r0=np.full(2 * n_mol, 1.0),
                  angle_atoms=angles, angle_k=np.full(n_mol, 50.0), angle_theta0=np.full(n_mol, 109.47),
                  kind="water")


def rocksalt_lattice(m: int, a: float = 1.0) -> System:
    """Perfect rock-salt lattice of m^3 sites (m even), spacing a, q = (-1)^(i+j+k)."""
    idx = np.arange(m)
    I, J, K = np.meshgrid(idx, idx, idx, indexing="ij")
    x = np.stack([I.ravel(), J.ravel(), K.ravel()], 1).astype(np.float64) * a
    q = np.where((I + J + K).ravel() % 2 == 0, 1.0, -1.0)
    n = m ** 3
    L = m * a
    return System(lo=np.zeros(3), hi=np.full(3, L), x=x, v=np.zeros((n, 3)),
                  type=np.zeros(n, np.int32), charge=q, mass=np.ones(n), eps=np.array([0.0]),
                  sigma=np.array([1.0]), kind="lattice")


def write_data(path: str, s: System, atom_style: str = "full", title: str = "rbmd-mi355x synthetic system"):
    """LAMMPS data file with exact box bounds (read back by read_data)."""
    ntypes = len(s.eps)
    lines = [title, ""]
    n = s.n_atoms
    nb = 0 if s.bond_atoms is None else len(s.bond_atoms) // 2
    na = 0 if s.angle_atoms is None else len(s.angle_atoms) // 3
    lines.append(f"{n} atoms")
    lines.append(f"{ntypes} atom types")
    if atom_style == "full":
        lines.append(f"{nb} bonds")
        lines.append(f"1 bond types")
        lines.append(f"{na} angles")
        lines.append(f"1 angle types")
    lines.append("")
    for d, nm in enumerate("xyz"):
        lines.append(f"{float(s.lo[d])!r} {float(s.hi[d])!r} {nm}lo {nm}hi")
    lines += ["", "Masses", ""]
    tmass = {}
    for i in range(n):
        tmass.setdefault(int(s.type[i]), float(s.mass[i]))
    for t in range(ntypes):
        lines.append(f"{t + 1} {tmass.get(t, 1.0)!r}")
    lines += ["", "Pair Coeffs", ""]
    for t in range(ntypes):
        lines.append(f"{t + 1} {float(s.eps[t])!r} {float(s.sigma[t])!r}")
    if atom_style == "full" and nb:
        lines += ["", "Bond Coeffs", "", f"1 {float(s.bond_k[0])!r} {float(s.bond_r0[0])!r}"]
        lines += ["", "Angle Coeffs", "", f"1 {float(s.angle_k[0])!r} {float(s.angle_theta0[0])!r}"]
    lines += ["", f"Atoms # {atom_style}", ""]
    for i in range(n):
        xs = " ".join(repr(float(c)) for c in s.x[i])
        if atom_style == "full":
            mol = int(s.mol_id[i]) + 1 if s.mol_id is not None else i + 1
            lines.append(f"{i + 1} {mol} {int(s.type[i]) + 1} {float(s.charge[i])!r} {xs}")
        elif atom_style == "charge":
            lines.append(f"{i + 1} {int(s.type[i]) + 1} {float(s.charge[i])!r} {xs}")
        else:
            lines.append(f"{i + 1} {int(s.type[i]) + 1} {xs}")
    lines += ["", "Velocities", ""]
    for i in range(n):
        lines.append(f"{i + 1} " + " ".join(repr(float(c)) for c in s.v[i]))
    if atom_style == "full" and nb:
        lines += ["", "Bonds", ""]
        for b in range(nb):
            lines.append(f"{b + 1} 1 {int(s.bond_atoms[2 * b]) + 1} {int(s.bond_atoms[2 * b + 1]) + 1}")
        lines += ["", "Angles", ""]
        for a in range(na):
            i0, i1, i2 = (int(s.angle_atoms[3 * a + k]) + 1 for k in range(3))
            lines.append(f"{a + 1} 1 {i0} {i1} {i2}")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def read_data(path: str, atom_style: str, unit: str, special_bonds=None) -> System:
    """LAMMPS data reader following ModelFileInitCondition (Header :120-208, Parser :210-278)."""
    with open(path, "r", encoding="utf-8-sig") as f:
        raw = f.read().splitlines()
    hdr = {}
    lo = np.zeros(3)
    hi = np.zeros(3)
    li = 0
    while li < len(raw):
        line = raw[li]
        li += 1
        if not line.strip():
            continue
        toks = line.split()
        for key in ("atoms", "bonds", "angles", "dihedrals", "impropers"):
            if line.find(key) != -1 and "types" not in line:
                hdr.setdefault(key, int(toks[0]))
                break
        if "atom types" in line:
            hdr["atom types"] = int(toks[0])
        elif "bond types" in line:
            hdr["bond types"] = int(toks[0])
        elif "angle types" in line:
            hdr["angle types"] = int(toks[0])
        elif "xlo xhi" in line:
            lo[0], hi[0] = float(toks[0]), float(toks[1])
        elif "ylo yhi" in line:
            lo[1], hi[1] = float(toks[0]), float(toks[1])
        elif "zlo zhi" in line:
            lo[2], hi[2] = float(toks[0]), float(toks[1])
            break
    n = hdr["atoms"]
    nt = hdr["atom types"]

    def take(count):
        nonlocal li
        out = []
        while len(out) < count and li < len(raw):
            line = raw[li]
            li += 1
            if line.strip() and line != "\r":
                out.append(line.split("#")[0].split())
        return out

    tmass = np.ones(nt)
    eps = np.zeros(nt)
    sig = np.zeros(nt)
    bk = br0 = ak = at0 = None
    x = np.zeros((n, 3))
    v = None
    typ = np.zeros(n, np.int32)
    q = np.zeros(n)
    mol = None
    bonds = angles = None
    btype = atype = None
    while li < len(raw):
        line = raw[li]
        li += 1
        if not line.strip():
            continue
        if "Masses" in line:
            for t in take(nt):
                tmass[int(float(t[0])) - 1] = float(t[1])
        elif "Pair Coeffs" in line:
            for t in take(nt):
                k = int(float(t[0])) - 1
                eps[k], sig[k] = float(t[1]), float(t[2])
        elif "Bond Coeffs" in line:
            rows = take(hdr["bond types"])
            bk = np.array([float(r[1]) for r in rows])
            br0 = np.array([float(r[2]) for r in rows])
        elif "Angle Coeffs" in line:
            rows = take(hdr["angle types"])
            ak = np.array([float(r[1]) for r in rows])
            at0 = np.array([float(r[2]) for r in rows])
        elif "Atoms" in line:
            rows = take(n)
            if atom_style == "full":
                mol = np.zeros(n, np.int32)
            for r in rows:
                i = int(r[0]) - 1
                if atom_style == "full":
                    mol[i] = int(r[1]) - 1
                    typ[i] = int(r[2]) - 1
                    q[i] = float(r[3])
                    x[i] = [float(r[4]), float(r[5]), float(r[6])]
                elif atom_style == "charge":
                    typ[i] = int(r[1]) - 1
                    q[i] = float(r[2])
                    x[i] = [float(r[3]), float(r[4]), float(r[5])]
                else:
                    typ[i] = int(r[1]) - 1
                    x[i] = [float(r[2]), float(r[3]), float(r[4])]
        elif "Bonds" in line:
            rows = take(hdr["bonds"])
            bonds = np.array([[int(r[2]) - 1, int(r[3]) - 1] for r in rows], np.int32).reshape(-1)
            btype = np.array([int(r[1]) - 1 for r in rows])
        elif "Angles" in line:
            rows = take(hdr["angles"])
            angles = np.array([[int(r[2]) - 1, int(r[3]) - 1, int(r[4]) - 1] for r in rows], np.int32).reshape(-1)
            atype = np.array([int(r[1]) - 1 for r in rows])
        elif "Velocities" in line:
            v = np.zeros((n, 3))
            for r in take(n):
                v[int(r[0]) - 1] = [float(r[1]), float(r[2]), float(r[3])]
    if unit == "LJ":
        tmass = np.ones(nt)  # ModelFileInitCondition.cpp:294-298
    s = System(lo=lo, hi=hi, x=x, v=v if v is not None else np.zeros((n, 3)), type=typ, charge=q,
               mass=tmass[typ], eps=eps, sigma=sig, mol_id=mol)
    if bonds is not None:
        s.bond_atoms = bonds
        s.bond_k = bk[btype]
        s.bond_r0 = br0[btype]
        if special_bonds is not None:
            s.special_offsets, s.special_ids, s.special_weights = special_lists(n, bonds, special_bonds)
    if angles is not None:
        s.angle_atoms = angles
        s.angle_k = ak[atype]
        s.angle_theta0 = at0[atype]
    s.velocities_read = v is not None
    return s
