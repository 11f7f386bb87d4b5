"""The LDS-tiled MIXED path at the sizes where it runs (at least two 2x2x2
cell blocks per compute unit): k_fast_force_tile on the 16-bit block lists,
its 32-bit fallback for blocks whose 4x4x4 staged cells hold more than 2048
atoms, the energy pass and U_cut over those lists. The reference for all of
it is the FP64 engine, whose per-pair arithmetic is the reference's and is
held to the oracle by test_gpu_fast.py; MIXED is held to the force-RMS bar
sqrt(sum|F - F64|^2 / sum|F64|^2) <= 1e-5."""
import numpy as np
import pytest

from moltwister_amd import abi, config
from moltwister_amd import workloads as W
from moltwister_amd.engine import Engine

pytestmark = pytest.mark.gpu

N = 131072            # 47.8 sigma box at rho* = 1.2: 512 to 729 cell blocks
FT_CAP = 2048


def rms_rel(F, Fr):
    return float(np.sqrt(np.sum((F - Fr) ** 2) / np.sum(Fr ** 2)))


def lj_cfg(engine, skin, ensemble=abi.NVE):
    c = config.defaults()
    c.engine = engine
    c.ensemble = ensemble
    c.cutoff_radius = 2.5
    c.neigh_shell = skin
    c.skin = skin
    c.timestep = 0.5
    c.temperature = 120.272
    c.temperature_relax = 20.0
    c.cutoff_force = 1000.0
    c.coulomb = abi.COUL_CUTOFF
    return c


def staged_totals(pos, L, rlist):
    """Atoms in the 4x4x4 staged cells of every 2x2x2 block of the FP64/MIXED
    cell grid (floor(L / rlist) cells per axis)."""
    nc = int(np.floor(L / rlist))
    c = np.floor((pos + L / 2.0) / (L / nc)).astype(np.int64) % nc
    cnt = np.zeros((nc, nc, nc), np.int64)
    np.add.at(cnt, (c[:, 0], c[:, 1], c[:, 2]), 1)
    for ax in range(3):
        cnt = sum(np.roll(cnt, -o, axis=ax) for o in (-1, 0, 1, 2))
    return cnt[::2, ::2, ::2].ravel()


@pytest.fixture(scope="module")
def mixture():
    return W.binary_lj_large(N, tiles=2, seed=5)


@pytest.mark.parametrize("skin", [0.3, 0.45])
def test_tile_forces_energies_ucut(mixture, skin):
    s = mixture
    tot = staged_totals(s.pos, s.box[0], 2.5 + skin)
    share16 = float(np.mean(tot <= FT_CAP))
    print(f"case tile forces skin={skin}: {tot.size} blocks, 16-bit share {share16:.2f}", flush=True)
    if skin == 0.3:
        assert share16 == 1.0
    else:
        assert 0.05 < share16 < 0.95          # both word formats in one system
    em = Engine(s, lj_cfg(abi.ENGINE_MIXED, skin))
    ef = Engine(s, lj_cfg(abi.ENGINE_FP64, skin))
    Fm, Um = em.compute_forces()            # the pass with per-atom energies
    Ff, Uf = ef.compute_forces()
    assert rms_rel(Fm, Ff) < 1e-5, rms_rel(Fm, Ff)
    scale = np.maximum(np.linalg.norm(Ff, axis=1), np.sqrt(np.mean(np.sum(Ff ** 2, 1))))
    err = np.linalg.norm(Fm - Ff, axis=1) / scale
    assert err.max() < 1e-3, err.max()
    assert abs(Um.sum(dtype=np.float64) - Uf.sum(dtype=np.float64)) <= 1e-5 * np.abs(Uf).sum(dtype=np.float64)
    # the force-only pass (the MD step's kernel) gives the energy pass's forces
    em.bench_kernel(0, 1)
    F0, _ = em.get_forces()
    assert np.abs(F0 - Fm).max() <= 1e-12 * np.abs(Fm).max(), np.abs(F0 - Fm).max()
    # U_cut: the same pairs inside rc by the same FP64 predicate
    tm, tf = em.thermo(), ef.thermo()
    assert tm["U_cut"] == pytest.approx(tf["U_cut"], rel=1e-12)
    # the virial sum F.r of the pressure: its terms are large and nearly
    # cancel, so the bound is relative to sum |F.r|
    fr_abs = float(np.sum(np.abs(np.sum(Ff * s.pos, axis=1))))
    assert abs(tm["sum_fr"] - tf["sum_fr"]) <= 1e-5 * fr_abs, (tm["sum_fr"], tf["sum_fr"], fr_abs)
    em.close()
    ef.close()


def test_tile_nve_tracks_fp64(mixture):
    """Descent (energy passes) then NVE with list rebuilds on the tile path
    follow the FP64 engine to the FP32 pair error."""
    print("case tile nve", flush=True)
    s = mixture
    relax = Engine(s, lj_cfg(abi.ENGINE_MIXED, 0.45))
    relax.optimize(40, 2e-5, 0.0)
    pos, mom = relax.get_state()
    relax.close()
    runs = []
    for eng in (abi.ENGINE_FP64, abi.ENGINE_MIXED):
        s2 = W.binary_lj_large(N, tiles=2, seed=5)
        s2.pos, s2.mom = pos, mom
        e = Engine(s2, lj_cfg(eng, 0.45))
        for _ in range(4):
            e.step(10)
        runs.append((e.thermo(), e.get_state(), e.stats()))
        e.close()
    (tf, (pf, _), sf), (tm, (pm, _), sm) = runs
    L = np.array(s.box)
    d = pm - pf
    d -= L * np.rint(d / L)
    assert np.abs(d).max() < 1e-3, np.abs(d).max()
    for k in ("T", "Etot", "U_cut"):
        assert tm[k] == pytest.approx(tf[k], rel=1e-4), (k, tm[k], tf[k])
    assert sm["rebuilds"] > 1 and sf["rebuilds"] > 1


def test_tile_cutoff_coulomb(mixture):
    """Charged mixture (random +-q, neutral) with cutoff Coulomb: the charges
    staged beside the positions."""
    print("case tile coulomb", flush=True)
    s = W.binary_lj_large(N, tiles=2, seed=5)
    q = np.full(N, 0.3)
    q[N // 2:] = -0.3
    np.random.default_rng(9).shuffle(q)
    s.charge = q
    em = Engine(s, lj_cfg(abi.ENGINE_MIXED, 0.3))
    ef = Engine(s, lj_cfg(abi.ENGINE_FP64, 0.3))
    Fm, Um = em.compute_forces()
    Ff, Uf = ef.compute_forces()
    assert rms_rel(Fm, Ff) < 1e-5, rms_rel(Fm, Ff)
    assert abs(Um.sum(dtype=np.float64) - Uf.sum(dtype=np.float64)) <= 1e-5 * np.abs(Uf).sum(dtype=np.float64)
    em.bench_kernel(0, 1)
    F0, _ = em.get_forces()
    assert np.abs(F0 - Fm).max() <= 1e-12 * np.abs(Fm).max()
    em.close()
    ef.close()
