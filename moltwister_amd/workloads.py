"""Seeded synthetic systems for the benchmark and parity workloads.

Reduced LJ units map onto the reference's internal units exactly
(Cmd/MolTwisterCmdMolDyn/Integrators/Constants.h:23-40): sigma = 1 Angstrom,
epsilon = 1 kJ/mol, m = 1 g/mol, so tau = 100 fs, T* = T / 120.272 K and
dt* = 0.005 is 0.5 fs.

Workloads:
  fcc_lj            monatomic LJ fluid on an FCC lattice (N = 4 n^3)
  water_box         flexible 3-site water, O-O LJ, harmonic O-H bonds and
                    H-O-H angles, point charges
  binary_lj         80:20 binary LJ (or Buckingham) mixture, random positions
                    with a minimum separation
  chain_molecules   short bonded chains with angles and dihedrals (exercises
                    every bonded term and the 1-2/1-3/1-4 scaling)
"""
from __future__ import annotations

import math

import numpy as np

import ctypes as C
import os

from . import abi
from .system import System

_GEN = None


def _genlib():
    global _GEN
    if _GEN is None:
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libmtgen.so")
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: python -m moltwister_amd.build")
        _GEN = C.CDLL(path)
        _GEN.mtgen_rsa_fast.argtypes = [C.c_int, C.c_double, C.c_double, C.c_uint64, C.c_int,
                                        C.POINTER(C.c_double)]
    return _GEN


def rsa_positions_c(n: int, L: float, dmin: float, seed: int) -> np.ndarray:
    """Random sequential addition in C (libmtgen.so, mtgen_rsa_fast)."""
    out = np.zeros((n, 3))
    k = _genlib().mtgen_rsa_fast(n, L, dmin, seed, 16, out.ctypes.data_as(C.POINTER(C.c_double)))
    if k != n:
        raise RuntimeError(f"rsa: placed {k} of {n}")
    return out

CONV_T = 120.272


def maxwell_boltzmann(mass: np.ndarray, T_kelvin: float, rng: np.random.Generator,
                      mobile: np.ndarray | None = None) -> np.ndarray:
    """Momenta ~ N(0, m kT), zero net momentum, rescaled so that the reference
    temperature sum p.p/m / (3N) (SimulationBox.cpp:168-180) equals T."""
    n = mass.shape[0]
    Tred = T_kelvin / CONV_T
    p = rng.standard_normal((n, 3)) * np.sqrt(mass * Tred)[:, None]
    mob = np.ones(n, bool) if mobile is None else mobile.astype(bool)
    if mob.sum() > 1:
        p[mob] -= p[mob].mean(axis=0)
    p[~mob] = 0.0
    Tnow = np.sum(np.sum(p * p, axis=1) / mass) / (3.0 * n)
    if Tnow > 0:
        p *= math.sqrt(Tred / Tnow)
    return p


def fcc_lj(ncell: int = 10, density: float = 0.8442, T_star: float = 0.72, seed: int = 1,
           eps: float = 1.0, sigma: float = 1.0) -> System:
    """Configuration 1: N = 4 ncell^3 LJ atoms on an FCC lattice centred on 0."""
    a = (4.0 / density) ** (1.0 / 3.0)
    L = ncell * a
    base = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]], float)
    g = np.arange(ncell, dtype=float)
    cells = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    pos = ((cells[:, None, :] + base[None, :, :]).reshape(-1, 3) + 0.25) * a - L / 2.0
    n = pos.shape[0]
    mass = np.ones(n)
    rng = np.random.default_rng(seed)
    mom = maxwell_boltzmann(mass, T_star * CONV_T, rng)
    return System(pos=pos, mass=mass, charge=np.zeros(n), types=np.zeros(n, np.int32),
                  box=(L, L, L), type_names=["Ar"], nonbonded=[(0, 0, abi.NB_LJ, (eps, sigma, 1.0))],
                  mom=mom, seed=seed)


def _random_rotations(rng: np.random.Generator, n: int) -> np.ndarray:
    """Uniform random rotation matrices from uniform unit quaternions."""
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    q = np.stack([np.sqrt(1 - u1) * np.sin(2 * np.pi * u2), np.sqrt(1 - u1) * np.cos(2 * np.pi * u2),
                  np.sqrt(u1) * np.sin(2 * np.pi * u3), np.sqrt(u1) * np.cos(2 * np.pi * u3)], -1)
    x, y, z, w = q.T
    R = np.empty((n, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z); R[:, 0, 1] = 2 * (x * y - z * w); R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w); R[:, 1, 1] = 1 - 2 * (x * x + z * z); R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w); R[:, 2, 1] = 2 * (y * z + x * w); R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


# SPC/Fw-like flexible water (literature values use U = k/2 d^2 in kcal/mol;
# the reference's harmonic terms are U = k d^2 in kJ/mol, MolTwisterMDFFBond_Harm.cpp:90-97,
# MolTwisterMDFFAngle_Harm.cpp:94-108, hence the factor 4.184/2)
WATER = dict(kb=1059.16 * 4.184 / 2.0, r0=1.012, ktheta=75.90 * 4.184 / 2.0, theta0=113.24,
             eps_OO=0.1554 * 4.184, sigma_OO=3.1655, qO=-0.82, qH=0.41, mO=15.9994, mH=1.008)


def water_box(n_mol: int = 10000, density_g_cm3: float = 1.0, T_kelvin: float = 298.0, seed: int = 2,
              jitter: float = 0.3, min_oo: float = 2.6, min_oh: float = 1.8, min_hh: float = 1.6,
              tries: int = 200) -> System:
    """Configuration 2 (and smaller boxes): waters on a jittered cubic grid at
    random orientations, atoms ordered O, H, H per molecule. Orientations are
    redrawn until every intermolecular O-O, O-H and H-H distance to the
    molecules already placed clears min_oo/min_oh/min_hh: the model has no LJ
    on H, so an H placed against a neighbouring O would collapse onto it
    under the bare Coulomb attraction."""
    w = WATER
    m_mol = w["mO"] + 2 * w["mH"]
    n_per_A3 = density_g_cm3 * 0.6022140857 / m_mol       # molecules per A^3
    L = (n_mol / n_per_A3) ** (1.0 / 3.0)
    g = int(math.ceil(n_mol ** (1.0 / 3.0)))
    rng = np.random.default_rng(seed)
    sites = np.stack(np.meshgrid(*(np.arange(g),) * 3, indexing="ij"), -1).reshape(-1, 3)
    sites = sites[rng.permutation(len(sites))[:n_mol]]
    spacing = L / g
    centres = (sites + 0.5) * spacing - L / 2.0 + rng.uniform(-jitter, jitter, (n_mol, 3))
    th = math.radians(w["theta0"])
    tmpl = np.array([[0.0, 0.0, 0.0],
                     [w["r0"], 0.0, 0.0],
                     [w["r0"] * math.cos(th), w["r0"] * math.sin(th), 0.0]])
    tmpl -= tmpl.mean(axis=0)
    # placed molecules by grid site, to find neighbours
    site_of = {tuple(sc): m for m, sc in enumerate(sites)}
    pos = np.zeros((n_mol, 3, 3))
    placed = np.zeros(n_mol, bool)
    is_h = np.array([False, True, True])
    thr = np.where(is_h[:, None] & is_h[None, :], min_hh,
                   np.where(~is_h[:, None] & ~is_h[None, :], min_oo, min_oh))
    offs = np.stack(np.meshgrid(*(np.arange(-1, 2),) * 3, indexing="ij"), -1).reshape(-1, 3)
    offs = offs[np.any(offs != 0, axis=1)]
    for m in range(n_mol):
        nb = [site_of.get(tuple((sites[m] + o) % g)) for o in offs]
        nb = [j for j in nb if j is not None and placed[j]]
        other = pos[nb] if nb else np.zeros((0, 3, 3))
        best, best_score = None, -np.inf
        for _ in range(tries):
            R = _random_rotations(rng, 1)[0]
            cand = centres[m] + tmpl @ R.T
            if len(nb) == 0:
                best = cand
                break
            d = cand[None, :, None, :] - other[:, None, :, :]
            d -= L * np.round(d / L)
            dist = np.sqrt(np.sum(d * d, axis=-1))            # (nb, 3 own, 3 other)
            score = np.min(dist - thr[None, :, :])
            if score > best_score:
                best, best_score = cand, score
            if score >= 0.0:
                break
        pos[m] = best
        placed[m] = True
    pos = pos.reshape(-1, 3)
    pos = (pos + L / 2.0) % L - L / 2.0
    n = 3 * n_mol
    types = np.tile(np.array([0, 1, 1], np.int32), n_mol)
    mass = np.tile(np.array([w["mO"], w["mH"], w["mH"]]), n_mol)
    charge = np.tile(np.array([w["qO"], w["qH"], w["qH"]]), n_mol)
    o = np.arange(n_mol) * 3
    bonds = np.concatenate([np.stack([o, o + 1, np.zeros(n_mol, int)], 1),
                            np.stack([o, o + 2, np.zeros(n_mol, int)], 1)], 0)
    bonds = bonds.reshape(2, n_mol, 3).transpose(1, 0, 2).reshape(-1, 3)
    angles = np.stack([o + 1, o, o + 2, np.zeros(n_mol, int)], 1)
    mom = maxwell_boltzmann(mass, T_kelvin, rng)
    return System(pos=pos, mass=mass, charge=charge, types=types, box=(L, L, L), type_names=["O", "H"],
                  nonbonded=[(0, 0, abi.NB_LJ, (w["eps_OO"], w["sigma_OO"], 1.0))],
                  bond_ff=[(abi.BOND_HARM, (w["kb"], w["r0"]))],
                  angle_ff=[(abi.ANGLE_HARM, (w["ktheta"], w["theta0"]))],
                  bonds=bonds, angles=angles, mom=mom, seed=seed)


def rsa_positions(n: int, L: float, dmin: float, rng: np.random.Generator, max_tries: int = 100000) -> np.ndarray:
    """Random sequential addition in a periodic cube with minimum separation
    dmin, accelerated by a cell grid (pure Python; for small n)."""
    nc = max(1, int(L // dmin))
    cs = L / nc
    grid: dict = {}
    out = np.empty((n, 3))
    d2 = dmin * dmin
    k = 0
    while k < n:
        for _ in range(max_tries):
            x = rng.uniform(-L / 2, L / 2, 3)
            c = np.floor((x + L / 2) / cs).astype(int) % nc
            ok = True
            for dx in (-1, 0, 1):
                for dy in (-1, 0, 1):
                    for dz in (-1, 0, 1):
                        for j in grid.get(((c[0] + dx) % nc, (c[1] + dy) % nc, (c[2] + dz) % nc), ()):
                            d = out[j] - x
                            d -= L * np.round(d / L)
                            if d @ d < d2:
                                ok = False
                                break
                        if not ok:
                            break
                    if not ok:
                        break
                if not ok:
                    break
            if ok:
                out[k] = x
                grid.setdefault(tuple(c), []).append(k)
                k += 1
                break
        else:
            raise RuntimeError("rsa_positions: packing too dense")
    return out


BINARY_LJ = {(0, 0): (1.0, 1.0), (0, 1): (1.5, 0.8), (1, 1): (0.5, 0.88)}


def binary_lj(n: int = 4000, density: float = 1.2, frac_a: float = 0.8, dmin: float = 0.8, seed: int = 3,
              T_star: float = 1.0, buckingham: bool = False, positions: np.ndarray | None = None) -> System:
    """Configuration 3 family: 80:20 A:B LJ mixture (eps_AA=1, eps_AB=1.5,
    eps_BB=0.5, sigma_AA=1, sigma_AB=0.8, sigma_BB=0.88) or the Buckingham
    variant (A=1000, rho=0.1, C=1 for all pairs). Types are in first-appearance
    order of a seeded shuffle."""
    L = (n / density) ** (1.0 / 3.0)
    rng = np.random.default_rng(seed)
    pos = rsa_positions(n, L, dmin, rng) if positions is None else positions
    na = int(round(frac_a * n))
    lab = np.array([0] * na + [1] * (n - na))
    lab = lab[rng.permutation(n)]
    # first-appearance order of the names
    first = lab[0]
    types = np.where(lab == first, 0, 1).astype(np.int32)
    names = ["A", "B"] if first == 0 else ["B", "A"]
    name_of = {0: names[0], 1: names[1]}
    if buckingham:
        nb = [(0, 0, abi.NB_BUCK, (1000.0, 0.1, 1.0)), (0, 1, abi.NB_BUCK, (1000.0, 0.1, 1.0)),
              (1, 1, abi.NB_BUCK, (1000.0, 0.1, 1.0))]
    else:
        idx = {name_of[t]: t for t in (0, 1)}
        nb = []
        for (a, b), (e, s) in BINARY_LJ.items():
            ta, tb = idx["AB"[a]], idx["AB"[b]]
            nb.append((ta, tb, abi.NB_LJ, (e, s, 1.0)))
    mass = np.ones(n)
    mom = maxwell_boltzmann(mass, T_star * CONV_T, rng)
    return System(pos=pos, mass=mass, charge=np.zeros(n), types=types, box=(L, L, L), type_names=names,
                  nonbonded=nb, mom=mom, seed=seed)


def binary_lj_large(n: int = 1048576, density: float = 1.2, tiles: int = 2, seed: int = 3, T_star: float = 1.0,
                    buckingham: bool = False) -> System:
    """Configuration 3 at full size: an RSA box of n/tiles^3 atoms (minimum
    separation 0.8 sigma, rho = 1.2) repeated tiles^3 times, then types and
    Maxwell-Boltzmann momenta drawn per atom, so the copies diverge from the
    first step. (Plain RSA of 10^6 points takes minutes near this packing;
    the periodic copy has period L/tiles = 47.8 sigma, far beyond the cutoff.)"""
    m = n // tiles ** 3
    if m * tiles ** 3 != n:
        raise ValueError("n must be divisible by tiles^3")
    L = (n / density) ** (1.0 / 3.0)
    base = rsa_positions_c(m, L / tiles, 0.8, seed)
    shifts = np.stack(np.meshgrid(*(np.arange(tiles),) * 3, indexing="ij"), -1).reshape(-1, 3)
    pos = (base[None, :, :] + (L / 2.0 / tiles) + shifts[:, None, :] * (L / tiles)).reshape(-1, 3) - L / 2.0
    return binary_lj(n, density=density, seed=seed, T_star=T_star, buckingham=buckingham, positions=pos)


def chain_molecules(n_chain: int = 60, length: int = 6, L: float = 24.0, seed: int = 4,
                    T_kelvin: float = 300.0) -> System:
    """Short bonded chains (C-C-...-C with alternating charges) at random
    positions and orientations, with harmonic bonds, angles and dihedrals and
    LJ + Coulomb between all atoms."""
    rng = np.random.default_rng(seed)
    b = 1.53
    th = math.radians(111.0)
    # planar zig-zag template then a random rotation per chain
    tmpl = np.zeros((length, 3))
    for i in range(1, length):
        tmpl[i] = tmpl[i - 1] + b * np.array([math.sin(th / 2), math.cos(th / 2) * (1 if i % 2 else -1), 0.0])
    tmpl -= tmpl.mean(axis=0)
    # random twist so dihedrals are not exactly planar
    tmpl[:, 2] += rng.normal(0, 0.2, length)
    centres = rsa_positions(n_chain, L, 6.5, rng)
    R = _random_rotations(rng, n_chain)
    pos = (centres[:, None, :] + np.einsum("mij,aj->mai", R, tmpl)).reshape(-1, 3)
    pos = (pos + L / 2) % L - L / 2
    n = n_chain * length
    types = np.tile(np.array([0 if i in (0, length - 1) else 1 for i in range(length)], np.int32), n_chain)
    mass = np.where(types == 0, 15.035, 14.027)
    charge = np.tile(np.array([0.1 if i % 2 == 0 else -0.1 for i in range(length)]), n_chain)
    bonds, angles, dih = [], [], []
    for c in range(n_chain):
        o = c * length
        for i in range(length - 1):
            bonds.append((o + i, o + i + 1, 0))
        for i in range(length - 2):
            angles.append((o + i, o + i + 1, o + i + 2, 0))
        for i in range(length - 3):
            dih.append((o + i, o + i + 1, o + i + 2, o + i + 3, 0))
    mom = maxwell_boltzmann(mass, T_kelvin, rng)
    return System(pos=pos, mass=mass, charge=charge, types=types, box=(L, L, L), type_names=["CH3", "CH2"],
                  nonbonded=[(0, 0, abi.NB_LJ, (0.8148, 3.75, 1.0)), (0, 1, abi.NB_LJ, (0.5, 3.85, 1.0)),
                             (1, 1, abi.NB_LJ, (0.3908, 3.95, 1.0))],
                  bond_ff=[(abi.BOND_HARM, (1000.0, 1.54))],
                  angle_ff=[(abi.ANGLE_HARM, (250.0, 114.0))],
                  dihedral_ff=[(abi.DIH_HARM, (5.0, 1, 3))],
                  bonds=np.array(bonds), angles=np.array(angles), dihedrals=np.array(dih), mom=mom, seed=seed)


def replica_lj(n_atoms: int, density: float, T_star: float, seed: int, jitter: float = 0.05) -> System:
    """One LJ replica (epsilon = sigma = m = 1): n_atoms on randomly chosen
    sites of an FCC lattice filling the cube of side (n/density)^(1/3),
    jittered by `jitter` lattice constants, Maxwell-Boltzmann momenta at T*."""
    rng = np.random.default_rng(seed)
    L = (n_atoms / density) ** (1.0 / 3.0)
    ncell = int(math.ceil((n_atoms / 4.0) ** (1.0 / 3.0) - 1e-9))
    a = L / ncell
    base = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]], float)
    g = np.arange(ncell, dtype=float)
    cells = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    sites = ((cells[:, None, :] + base[None, :, :]).reshape(-1, 3) + 0.25) * a - L / 2.0
    pick = np.sort(rng.choice(sites.shape[0], n_atoms, replace=False))
    pos = sites[pick] + rng.uniform(-jitter * a, jitter * a, (n_atoms, 3))
    mass = np.ones(n_atoms)
    mom = maxwell_boltzmann(mass, T_star * CONV_T, rng)
    return System(pos=pos, mass=mass, charge=np.zeros(n_atoms), types=np.zeros(n_atoms, np.int32),
                  box=(L, L, L), type_names=["Ar"], nonbonded=[(0, 0, abi.NB_LJ, (1.0, 1.0, 1.0))],
                  mom=mom, seed=seed)


def replicas(n: int = 512, seed: int = 7, n_min: int = 256, n_max: int = 4096):
    """Configuration 4: n independent LJ systems with sizes uniform in
    [n_min, n_max], densities uniform in [0.6, 1.0] and temperatures T*
    uniform in [0.6, 1.5] (seeded). Returns (systems, T_star list)."""
    rng = np.random.default_rng(seed)
    sizes = rng.integers(n_min, n_max + 1, n)
    dens = rng.uniform(0.6, 1.0, n)
    temps = rng.uniform(0.6, 1.5, n)
    systems = [replica_lj(int(sizes[i]), float(dens[i]), float(temps[i]), seed * 100003 + i) for i in range(n)]
    return systems, [float(t) for t in temps]


def water_frames(n_mol: int, n_frames: int, every: int, seed: int = 2, T_kelvin: float = 298.0,
                 cutoff: float = 10.0, skin: float = 0.3, relax_steps: int = 300, equil_steps: int = 200,
                 progress=None):
    """Configuration 5 frames: a flexible-water box (configuration 2 force
    field) relaxed by gradient descent, then Nose-Hoover NVT with all-pairs
    Coulomb on the MIXED engine, sampled every `every` steps after
    `equil_steps`. Returns (system, frames[n_frames, 3 n_atoms])."""
    from .engine import Engine
    from . import config
    s = water_box(n_mol, seed=seed)
    c = config.defaults()
    c.engine, c.coulomb = abi.ENGINE_MIXED, abi.COUL_ALLPAIRS
    c.cutoff_radius, c.neigh_shell, c.skin = cutoff, skin, skin
    c.timestep, c.temperature, c.temperature_relax, c.ensemble = 0.5, T_kelvin, 20.0, abi.NVT
    e = Engine(s, c)
    e.optimize(relax_steps, 1e-4, 0.0)
    e.step(equil_steps)
    frames = np.zeros((n_frames, s.n_atoms * 3))
    for f in range(n_frames):
        e.step(every)
        frames[f] = e.get_state()[0].reshape(-1)
        if progress and (f + 1) % 128 == 0:
            progress(f + 1)
    e.close()
    return s, frames
