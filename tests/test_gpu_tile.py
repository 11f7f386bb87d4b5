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
    # the measurement pass with FP64 arithmetic on the same lists: the FP64
    # engine's forces up to the FP32 segment coefficients a_f + b_f / r that
    # both tiled passes use (about 1e-6 against the reference's interpolation),
    # below MIXED's error; 32-bit blocks refused
    if share16 == 1.0:
        em.bench_kernel(4, 1)
        F64t, _ = em.get_forces()
        assert rms_rel(F64t, Ff) < 2e-6 and rms_rel(F64t, Ff) < rms_rel(Fm, Ff), (rms_rel(F64t, Ff), rms_rel(Fm, Ff))
        em.bench_kernel(0, 1)
    else:
        with pytest.raises(Exception, match="32-bit"):
            em.bench_kernel(4, 1)
        em.bench_kernel(0, 1)
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


@pytest.mark.parametrize("coul", [abi.COUL_CUTOFF, abi.COUL_ALLPAIRS])
def test_tile_coulomb(mixture, coul):
    """Charged mixture (random +-q, neutral): cutoff Coulomb with the charges
    staged beside the positions, and the reference's all-pairs Coulomb (its
    own N^2 pass beside the tile kernel)."""
    print(f"case tile coulomb {coul}", flush=True)
    s = W.binary_lj_large(N, tiles=2, seed=5)
    q = np.full(N, 0.3)
    q[N // 2:] = -0.3
    np.random.default_rng(9).shuffle(q)
    s.charge = q
    cm, cf = lj_cfg(abi.ENGINE_MIXED, 0.3), lj_cfg(abi.ENGINE_FP64, 0.3)
    cm.coulomb = cf.coulomb = coul
    em = Engine(s, cm)
    ef = Engine(s, cf)
    Fm, Um = em.compute_forces()
    Ff, Uf = ef.compute_forces()
    assert rms_rel(Fm, Ff) < 1e-5, rms_rel(Fm, Ff)
    assert abs(Um.sum(dtype=np.float64) - Uf.sum(dtype=np.float64)) <= 1e-5 * np.abs(Uf).sum(dtype=np.float64)
    em.bench_kernel(0, 1)
    F0, _ = em.get_forces()
    assert np.abs(F0 - Fm).max() <= 1e-12 * np.abs(Fm).max()
    em.close()
    ef.close()


def water_cfg(engine, ensemble=abi.NVT):
    c = config.defaults()
    c.engine = engine
    c.ensemble = ensemble
    c.cutoff_radius = 5.5
    c.neigh_shell = 1.0
    c.skin = 1.0
    c.timestep = 0.5
    c.temperature = 298.0
    c.temperature_relax = 20.0
    c.coulomb = abi.COUL_CUTOFF
    return c


def test_tile_water_molecules():
    """Water (bonds and angles: same-molecule pairs left to the special-pair
    pass) on the tile path: 114,000 atoms in a 104 A box, 16 cells of 6.5 A a
    side, 512 blocks, every block on 16-bit words."""
    print("case tile water", flush=True)
    s = W.water_box(38000, seed=3)
    tot = staged_totals(s.pos, s.box[0], 6.5)
    assert tot.size == 512 and tot.max() <= FT_CAP
    em, ef = Engine(s, water_cfg(abi.ENGINE_MIXED)), Engine(s, water_cfg(abi.ENGINE_FP64))
    Fm, Um = em.compute_forces()
    Ff, Uf = ef.compute_forces()
    assert rms_rel(Fm, Ff) < 1e-5, rms_rel(Fm, Ff)
    assert abs(Um.sum(dtype=np.float64) - Uf.sum(dtype=np.float64)) <= 1e-5 * np.abs(Uf).sum(dtype=np.float64)
    em.bench_kernel(0, 1)
    F0, _ = em.get_forces()
    assert np.abs(F0 - Fm).max() <= 1e-12 * np.abs(Fm).max()
    assert em.thermo()["U_cut"] == pytest.approx(ef.thermo()["U_cut"], rel=1e-12)
    for e in (em, ef):
        e.step(20)
    tm, tf = em.thermo(), ef.thermo()
    for k in ("T", "Etot"):
        assert tm[k] == pytest.approx(tf[k], rel=1e-4), (k, tm[k], tf[k])
    pm, _ = em.get_state()
    pf, _ = ef.get_state()
    L = np.array(s.box)
    d = pm - pf
    d -= L * np.rint(d / L)
    assert np.abs(d).max() < 1e-3, np.abs(d).max()
    em.close()
    ef.close()


def test_tile_npt_follows_fp64(mixture):
    """NPT on the tile path: the staged block frame follows the cell width of
    the current box every step, and a regrid (the box shrinking below 17 cells
    of cutoff + skin) switches blocks between word formats; the run tracks the
    FP64 engine. The barostat mass W = T tau_P^2 has no factor N, so a
    131,072-atom box needs a long tau_P; it starts with an inward p_eps at the
    current pressure."""
    print("case tile npt", flush=True)
    relax = Engine(mixture, lj_cfg(abi.ENGINE_MIXED, 0.3))
    relax.optimize(40, 2e-5, 0.0)
    pos, mom = relax.get_state()
    P0 = relax.thermo()["P"]
    relax.close()
    T = 120.272 / 120.272
    tauP = 50000.0 / 100.0
    Wb = T * tauP * tauP
    runs = []
    for eng in (abi.ENGINE_FP64, abi.ENGINE_MIXED):
        s2 = W.binary_lj_large(N, tiles=2, seed=5)
        s2.pos, s2.mom = pos, mom
        c = lj_cfg(eng, 0.3, ensemble=abi.NPT)
        c.pressure = P0
        c.pressure_relax = 50000.0
        c.temperature_relax = 50.0
        c.cutoff_force = 0.0
        e = Engine(s2, c)
        e.set_barostat(0.0, -15.0 * Wb / tauP, None, None)
        trace = []
        for _ in range(4):
            e.step(25)
            th, b = e.thermo(), e.get_barostat()
            trace.append((th["V"], th["T"], b["box"][0]))
        print(f"engine {eng}: P0 {P0:.1f}, (V, T, L) {trace}", flush=True)
        runs.append((trace, e.get_barostat(), e.stats()))
        e.close()
    (trf, bf, sf), (trm, bm, sm) = runs
    for (vf, tf_, lf), (vm, tm_, lm) in zip(trf, trm):
        assert vm == pytest.approx(vf, rel=3e-5)
        assert tm_ == pytest.approx(tf_, rel=1e-3)
    # the box moved through a regrid: below 17 cells of 2.8 sigma (47.6 sigma)
    assert min(lf for _, _, lf in trf) < 17 * 2.8
    assert abs(bf["eps"]) > 1e-4
    assert np.allclose(bm["box"], bf["box"], rtol=1e-5, atol=0)
    assert sm["rebuilds"] > 1


def test_tile_monatomic_fcc():
    """One atom type (T = 1): a 256,000-atom FCC LJ crystal (configuration 1's
    lattice at 40^3 unit cells, 1728 cell blocks), positions perturbed off the
    lattice sites, forces and energies against FP64."""
    print("case tile fcc", flush=True)
    s = W.fcc_lj(40)
    s.pos = s.pos + np.random.default_rng(8).normal(0.0, 0.03, s.pos.shape)
    em, ef = Engine(s, lj_cfg(abi.ENGINE_MIXED, 0.3)), Engine(s, lj_cfg(abi.ENGINE_FP64, 0.3))
    Fm, Um = em.compute_forces()
    Ff, Uf = ef.compute_forces()
    assert rms_rel(Fm, Ff) < 1e-5, rms_rel(Fm, Ff)
    assert abs(Um.sum(dtype=np.float64) - Uf.sum(dtype=np.float64)) <= 1e-5 * np.abs(Uf).sum(dtype=np.float64)
    em.bench_kernel(0, 1)
    F0, _ = em.get_forces()
    assert np.abs(F0 - Fm).max() <= 1e-12 * np.abs(Fm).max()
    assert em.thermo()["U_cut"] == pytest.approx(ef.thermo()["U_cut"], rel=1e-12)
    em.close()
    ef.close()
