"""GPU checks of the FP64 and MIXED engines (skin lists, sorted atoms).

FP64 runs the reference's per-pair arithmetic on a different summation order,
so its forces match the oracle to 1e-12 relative per atom. MIXED computes pair
geometry and interpolation in FP32 and is held to the force-RMS bar
sqrt(sum|F - F_ref|^2 / sum|F_ref|^2) <= 1e-5 against the FP64 oracle.
Coulomb is the cutoff variant in both (the oracle's COUL_CUTOFF mode).
"""
import numpy as np
import pytest

from moltwister_amd import abi
from moltwister_amd.engine import Engine

from systems import make

pytestmark = pytest.mark.gpu

FAST = ["binary_lj", "buckingham", "water_300", "fcc_4000", "chains_class2"]


def rms_rel(F, Fr):
    return float(np.sqrt(np.sum((F - Fr) ** 2) / np.sum(Fr ** 2)))


@pytest.mark.parametrize("name", FAST)
@pytest.mark.parametrize("engine", [abi.ENGINE_FP64, abi.ENGINE_MIXED])
def test_forces_vs_oracle(oracle_mod, name, engine):
    print(f"case fast forces {name} engine={engine}", flush=True)
    s, c = make(name, engine=engine, coulomb=abi.COUL_CUTOFF)
    co = make(name, coulomb=abi.COUL_CUTOFF)[1]
    o, e = oracle_mod.Oracle(s, co), Engine(s, c)
    # move atoms off lattice sites: a perfect lattice has forces at rounding level
    pos = s.pos + np.random.default_rng(3).normal(0, 0.03, s.pos.shape)
    o.set_state(pos=pos)
    e.set_state(pos=pos)
    Fo, Uo = o.compute_forces()
    Fe, Ue = e.compute_forces()
    if engine == abi.ENGINE_FP64:
        scale = np.maximum(np.linalg.norm(Fo, axis=1), np.sqrt(np.mean(np.sum(Fo ** 2, 1))))
        err = np.linalg.norm(Fe - Fo, axis=1) / scale
        assert err.max() < 1e-12, err.max()
        assert abs(Ue.astype(np.float64).sum() - Uo.astype(np.float64).sum()) <= 1e-5 * np.abs(Uo).sum()
    else:
        assert rms_rel(Fe, Fo) < 1e-5, rms_rel(Fe, Fo)
        assert abs(Ue.astype(np.float64).sum() - Uo.astype(np.float64).sum()) <= 1e-4 * np.abs(Uo).sum()


@pytest.mark.parametrize("engine", [abi.ENGINE_FP64, abi.ENGINE_MIXED])
def test_trajectory_vs_exact(oracle_mod, engine):
    """50 NVE steps of the LJ fluid: skin lists with rebuilds and re-sorts give
    the same trajectory as the reference algorithm to rounding (FP64) or to
    the FP32 pair error (MIXED)."""
    print(f"case fast trajectory engine={engine}", flush=True)
    s, c = make("fcc_4000", engine=engine)
    ce = make("fcc_4000")[1]
    ref, e = Engine(s, ce), Engine(s, c)
    for _ in range(5):
        ref.step(10)
        e.step(10)
    pr, mr = ref.get_state()
    pe, me = e.get_state()
    tol = 1e-9 if engine == abi.ENGINE_FP64 else 2e-4
    assert np.abs(pe - pr).max() < tol, np.abs(pe - pr).max()
    tr, te = ref.thermo(), e.thermo()
    assert abs(te["Etot"] - tr["Etot"]) < (1e-9 if engine == abi.ENGINE_FP64 else 1e-4) * abs(tr["Etot"])
    st = e.stats()
    assert st["rebuilds"] >= 1


def test_state_roundtrip_sorted(oracle_mod):
    """get_state/set_state return the caller's atom order whatever the
    engine's internal (spatially sorted) order is."""
    s, c = make("binary_lj", engine=abi.ENGINE_MIXED)
    e = Engine(s, c)
    e.step(30)
    pos, mom = e.get_state()
    e.set_state(pos=pos, mom=mom)
    p2, m2 = e.get_state()
    assert np.array_equal(pos, p2) and np.array_equal(mom, m2)
    F1, _ = e.compute_forces()
    o = oracle_mod.Oracle(s, make("binary_lj")[1])
    o.set_state(pos=pos)
    Fo, _ = o.compute_forces()
    assert rms_rel(F1, Fo) < 1e-5


@pytest.mark.parametrize("engine", [abi.ENGINE_FP64, abi.ENGINE_MIXED])
def test_nvt_matches_exact_engine(engine):
    """NVT (Nose-Hoover chain on the device) on a relaxed 16,384-atom mixture:
    20 steps of the sorted skin-list engines track the reference-algorithm
    engine (FP64 to rounding, MIXED to its FP32 pair error)."""
    from moltwister_amd import workloads as W
    print(f"case nvt 16384 engine={engine}", flush=True)
    s = W.binary_lj_large(16384, tiles=1, seed=3)
    base = make("binary_lj", engine=abi.ENGINE_EXACT, ensemble=abi.NVE, T=120.272, dt=0.5)[1]
    relax = Engine(s, base)
    relax.optimize(100, 2e-5, 0.0)
    relax.step(100)
    pos, mom = relax.get_state()
    T = relax.thermo()["T"]
    s.pos, s.mom = pos, mom * np.sqrt(120.272 / T)
    runs = []
    for eng in (abi.ENGINE_EXACT, engine):
        c = make("binary_lj", engine=eng, ensemble=abi.NVT, T=120.272, dt=0.5)[1]
        e = Engine(s, c)
        e.step(20)
        runs.append((e.thermo(), e.get_state(), e.get_thermostat()))
    (t0, (p0, m0), nh0), (t1, (p1, m1), nh1) = runs
    tol = 1e-9 if engine == abi.ENGINE_FP64 else 1e-4
    assert np.abs(p1 - p0).max() < tol * 10, np.abs(p1 - p0).max()
    assert abs(t1["T"] - t0["T"]) <= tol * abs(t0["T"]), (t1["T"], t0["T"])
    # per-atom energies are float sums in both engines (the reference keeps U_k
    # as float); FP64 splits each pair's energy evenly, the reference gives it
    # to the atom with the lower index, so totals agree to float rounding
    assert abs(t1["Etot"] - t0["Etot"]) <= max(tol, 1e-7) * abs(t0["Etot"]), (t1["Etot"], t0["Etot"])
    assert np.allclose(nh1[1], nh0[1], rtol=tol * 10, atol=tol)


@pytest.mark.parametrize("engine", [abi.ENGINE_EXACT, abi.ENGINE_MIXED])
def test_nonfinite_position_reported(engine):
    """A NaN coordinate is reported as MTMD_ENAN (no list is built from it)."""
    from moltwister_amd.engine import MtmdError
    print(f"case nan engine={engine}", flush=True)
    s, c = make("binary_lj", engine=engine)
    e = Engine(s, c)
    pos = s.pos.copy()
    pos[17, 1] = np.nan
    e.set_state(pos=pos)
    with pytest.raises(MtmdError) as ex:
        e.compute_forces()
    assert ex.value.status == abi.MTMD_ENAN
    e.set_state(pos=s.pos)
    F, _ = e.compute_forces()
    assert np.isfinite(F).all()


@pytest.mark.parametrize("name", ["water_300", "chains_rc7"])
@pytest.mark.parametrize("engine", [abi.ENGINE_FP64, abi.ENGINE_MIXED])
def test_allpairs_coulomb_vs_oracle(oracle_mod, name, engine):
    """All-pairs minimum-image Coulomb (the reference's own Coulomb) in the
    sorted engines against the oracle's all-pairs forces."""
    print(f"case allpairs {name} engine={engine}", flush=True)
    s, c = make(name, engine=engine, coulomb=abi.COUL_ALLPAIRS)
    co = make(name, coulomb=abi.COUL_ALLPAIRS)[1]
    o, e = oracle_mod.Oracle(s, co), Engine(s, c)
    pos = s.pos + np.random.default_rng(4).normal(0, 0.02, s.pos.shape)
    o.set_state(pos=pos)
    e.set_state(pos=pos)
    Fo, Uo = o.compute_forces()
    Fe, Ue = e.compute_forces()
    if engine == abi.ENGINE_FP64:
        scale = np.maximum(np.linalg.norm(Fo, axis=1), np.sqrt(np.mean(np.sum(Fo ** 2, 1))))
        assert (np.linalg.norm(Fe - Fo, axis=1) / scale).max() < 1e-11
        assert abs(Ue.astype(np.float64).sum() - Uo.astype(np.float64).sum()) <= 1e-5 * np.abs(Uo).sum()
    else:
        assert rms_rel(Fe, Fo) < 1e-5, rms_rel(Fe, Fo)
        assert abs(Ue.astype(np.float64).sum() - Uo.astype(np.float64).sum()) <= 1e-4 * np.abs(Uo).sum()


@pytest.mark.parametrize("engine", [abi.ENGINE_FP64, abi.ENGINE_MIXED])
def test_water_nve_energy_allpairs(engine):
    """Relaxed water with all-pairs Coulomb: 200 NVE steps conserve the total
    energy (bonded, angle, LJ and Coulomb terms through the sorted engines)."""
    from moltwister_amd import workloads as W
    print(f"case water nve engine={engine}", flush=True)
    s = W.water_box(300, seed=5)
    c = make("water_300", engine=engine, coulomb=abi.COUL_ALLPAIRS)[1]
    c.cutoff_force = 1e9
    e = Engine(s, c)
    e.optimize(300, 1e-4, 0.0)
    e.step(20)
    E0 = e.thermo()["Etot"]
    e.step(200)
    th = e.thermo()
    assert np.isfinite(th["Etot"])
    assert abs(th["Etot"] - E0) < 2e-3 * abs(E0), (th["Etot"], E0)


def test_ucut_brute_force_and_engines_agree():
    """U_cut (pairs inside rc times the pair energy at rc) from all three
    engines against an O(N^2) numpy count with the same float-box predicate."""
    from moltwister_amd.engine import profile
    print("case ucut binary_lj", flush=True)
    s, c = make("binary_lj", engine=abi.ENGINE_EXACT)
    pos = s.pos + np.random.default_rng(8).normal(0, 0.02, s.pos.shape)
    L = np.array(s.box)
    d = pos[:, None, :] - pos[None, :, :]
    d -= L * np.round(d / L)
    r2 = np.sum(d * d, axis=2)
    inside = (r2 < float(np.float32(c.cutoff_radius)) ** 2) & ~np.eye(s.n_atoms, dtype=bool)
    ti = s.types
    T = max(s.n_types, max(max(a, b) + 1 for a, b, _, _ in s.nonbonded))
    ecut = np.zeros((T, T))
    for a, b, kind, prm in s.nonbonded:
        _, _, e = profile(0, kind, prm, 0.01, float(np.float32(c.cutoff_radius)))
        ecut[a, b] += float(e[-1])
        if a != b:
            ecut[b, a] += float(e[-1])
    want = 0.5 * float(np.sum(ecut[ti[:, None], ti[None, :]] * inside))
    got = {}
    for eng in (abi.ENGINE_EXACT, abi.ENGINE_FP64, abi.ENGINE_MIXED):
        e = Engine(s, make("binary_lj", engine=eng)[1])
        e.set_state(pos=pos)
        e.compute_forces()
        got[eng] = e.thermo()["U_cut"]
    for eng, v in got.items():
        assert abs(v - want) <= 1e-9 * abs(want), (eng, v, want)


@pytest.mark.parametrize("engine", [abi.ENGINE_EXACT, abi.ENGINE_FP64, abi.ENGINE_MIXED])
def test_ucut_cutoff_coulomb_water(engine):
    """U_cut with Coulomb truncated at rc: every pair of different molecules
    inside rc adds its tabulated U(rc) and K qi qj / rc (FP32, as the pair
    term forms it); pairs inside a water (1-2, 1-3) are scaled by 0. Against
    an O(N^2) numpy sum with the same float-box predicate."""
    from moltwister_amd.engine import profile
    print(f"case ucut water cutoff coulomb engine={engine}", flush=True)
    s, c = make("water_300", engine=engine, coulomb=abi.COUL_CUTOFF)
    pos = s.pos + np.random.default_rng(9).normal(0, 0.02, s.pos.shape)
    L = np.array(s.box)
    d = pos[:, None, :] - pos[None, :, :]
    d -= L * np.round(d / L)
    r2 = np.sum(d * d, axis=2)
    rcf = np.float32(c.cutoff_radius)
    mol = np.arange(s.n_atoms) // 3
    inside = (r2 < float(rcf) ** 2) & (mol[:, None] != mol[None, :])
    ti = s.types
    T = s.n_types
    ecut = np.zeros((T, T), np.float32)
    for a, b, kind, prm in s.nonbonded:
        _, _, e = profile(0, kind, prm, 0.01, float(rcf))
        ecut[a, b] += e[-1]
        if a != b:
            ecut[b, a] += e[-1]
    q = s.charge.astype(np.float32)
    ec = (np.float32(1389.32) * q[:, None] * q[None, :]) / rcf
    terms = (ecut[ti[:, None], ti[None, :]] + ec).astype(np.float64) * inside
    want = 0.5 * float(np.sum(terms))
    e = Engine(s, c)
    e.set_state(pos=pos)
    e.compute_forces()
    got = e.thermo()["U_cut"]
    e.close()
    assert abs(got - want) <= 1e-9 * float(np.sum(np.abs(terms))), (got, want)


@pytest.mark.parametrize("engine", [abi.ENGINE_EXACT, abi.ENGINE_FP64, abi.ENGINE_MIXED])
def test_profile_kind_matches_parameters(engine):
    """MTMD_FF_PROFILE force fields give bit-identical forces to the same force
    fields given by parameters, in every engine."""
    from test_oracle_golden import as_profiles
    print(f"case profile kind engine={engine}", flush=True)
    s, c = make("water_300", engine=engine, coulomb=abi.COUL_ALLPAIRS)
    a = Engine(s, c).compute_forces()
    b = Engine(as_profiles(s, c.cutoff_radius), c).compute_forces()
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


@pytest.mark.parametrize("coul", [abi.COUL_ALLPAIRS, abi.COUL_CUTOFF])
@pytest.mark.parametrize("engine", [abi.ENGINE_FP64, abi.ENGINE_MIXED])
def test_dense_cells_water_1000(oracle_mod, engine, coul):
    """1000 waters with a 10 A cutoff: ~110 atoms per cell, the dense-cell
    list builder (one workgroup per home cell)."""
    from moltwister_amd import workloads as W
    from systems import cfg
    print(f"case dense water engine={engine} coul={coul}", flush=True)
    s = W.water_box(1000, seed=6)
    c = cfg(10.0, 0.3, engine=engine, coulomb=coul, skin=0.3)
    co = cfg(10.0, 0.3, coulomb=coul)
    pos = s.pos + np.random.default_rng(7).normal(0, 0.02, s.pos.shape)
    o, e = oracle_mod.Oracle(s, co), Engine(s, c)
    o.set_state(pos=pos)
    e.set_state(pos=pos)
    Fo, Uo = o.compute_forces()
    Fe, Ue = e.compute_forces()
    if engine == abi.ENGINE_FP64:
        scale = np.maximum(np.linalg.norm(Fo, axis=1), np.sqrt(np.mean(np.sum(Fo ** 2, 1))))
        assert (np.linalg.norm(Fe - Fo, axis=1) / scale).max() < 1e-11
    else:
        assert rms_rel(Fe, Fo) < 1e-5, rms_rel(Fe, Fo)
    assert abs(Ue.astype(np.float64).sum() - Uo.astype(np.float64).sum()) <= 1e-4 * np.abs(Uo).sum()


@pytest.mark.parametrize("engine", [abi.ENGINE_FP64, abi.ENGINE_MIXED])
def test_fused_nve_steps_equal_single_steps(engine):
    """NVE steps inside one call run the back half of a step and the front half
    of the next in one kernel; the trajectory is bit-identical to the same
    steps taken one call at a time (separate kernels)."""
    print(f"case fused nve engine={engine}", flush=True)
    s, c = make("binary_lj", engine=engine)
    s.pos = s.pos + np.random.default_rng(12).normal(0, 0.01, s.pos.shape)
    a, b = Engine(s, c), Engine(s, c)
    a.step(60)
    for _ in range(60):
        b.step(1)
    pa, ma = a.get_state()
    pb, mb = b.get_state()
    assert np.array_equal(pa, pb) and np.array_equal(ma, mb)
    assert a.thermo()["Etot"] == b.thermo()["Etot"]


@pytest.mark.parametrize("engine", [abi.ENGINE_FP64, abi.ENGINE_MIXED])
def test_graph_replayed_nvt_steps_equal_single_steps(engine):
    """Plain NVT steps inside one call replay a captured step graph (re-sorts
    every 100 steps swap the buffers and re-capture); the trajectory equals
    the same steps taken one call at a time, bit for bit."""
    print(f"case graph nvt engine={engine}", flush=True)
    s, c = make("fcc_4000", engine=engine, ensemble=abi.NVT, T=86.596)
    s.pos = s.pos + np.random.default_rng(13).normal(0, 0.01, s.pos.shape)
    a, b = Engine(s, c), Engine(s, c)
    a.step(230)
    for _ in range(230):
        b.step(1)
    pa, ma = a.get_state()
    pb, mb = b.get_state()
    assert np.array_equal(pa, pb) and np.array_equal(ma, mb)
    ea, pea = a.get_thermostat()
    eb, peb = b.get_thermostat()
    assert np.array_equal(ea, eb) and np.array_equal(pea, peb)


@pytest.mark.parametrize("rc,per_cell", [(6.0, (64, 512)), (9.0, (512, 2048))])
@pytest.mark.parametrize("engine", [abi.ENGINE_FP64, abi.ENGINE_MIXED])
def test_dense_lj_cells_sorted_either_way(oracle_mod, engine, rc, per_cell):
    """An LJ fluid of 19,652 atoms with a long cutoff: at rc = 6 the cells hold
    a few hundred atoms (the cell sort ranks them in chunks of whole cells in
    LDS), at rc = 9 more than 512 (a cell larger than a chunk, sorted by one
    lane). Forces against the oracle, and two engines built from the same
    state give the same bits (the cells' order is deterministic)."""
    from moltwister_amd import workloads as W
    from systems import cfg
    print(f"case dense lj cells engine={engine} rc={rc}", flush=True)
    s = W.fcc_lj(17)
    L = s.box[0]
    nc = int(np.floor(L / (rc + 0.3)))
    assert nc >= 3 and per_cell[0] < s.n_atoms / nc ** 3 <= per_cell[1], (nc, s.n_atoms / nc ** 3)
    c = cfg(rc, 0.3, engine=engine, coulomb=abi.COUL_CUTOFF, skin=0.3)
    co = cfg(rc, 0.3, coulomb=abi.COUL_CUTOFF)
    pos = s.pos + np.random.default_rng(8).normal(0, 0.05, s.pos.shape)
    s.pos = pos
    o = oracle_mod.Oracle(s, co)
    Fo, Uo = o.compute_forces()
    del o
    e1 = Engine(s, c)
    Fe, Ue = e1.compute_forces()
    e1.close()
    e2 = Engine(s, c)
    Fe2, Ue2 = e2.compute_forces()
    e2.close()
    assert np.array_equal(Fe, Fe2) and np.array_equal(Ue, Ue2)
    if engine == abi.ENGINE_FP64:
        scale = np.maximum(np.linalg.norm(Fo, axis=1), np.sqrt(np.mean(np.sum(Fo ** 2, 1))))
        assert (np.linalg.norm(Fe - Fo, axis=1) / scale).max() < 1e-12
    else:
        assert rms_rel(Fe, Fo) < 1e-5, rms_rel(Fe, Fo)
    assert abs(Ue.astype(np.float64).sum() - Uo.astype(np.float64).sum()) <= 1e-5 * np.abs(Uo).sum()


@pytest.mark.parametrize("engine", [abi.ENGINE_FP64, abi.ENGINE_MIXED])
def test_crowded_cell_falls_back_to_dense_build(oracle_mod, engine):
    """A liquid-density FCC cube of 2916 atoms inside one cell of a 75 sigma
    box (rc 14.5, 5^3 cells): the mean cell is sparse, so the block list build
    runs first and meets a cell larger than its staging; the engine moves to
    the dense-cell build (any cell size) and goes on. Forces against the
    oracle, whose cells are sized by their counts as the reference's are
    (FunctorGenCellList.cpp:95-137), and steps that keep the energy."""
    from moltwister_amd.system import System
    from systems import cfg
    print(f"case crowded cell engine={engine}", flush=True)
    a = (4.0 / 1.0) ** (1.0 / 3.0)
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    g = np.stack(np.meshgrid(*[np.arange(9)] * 3, indexing="ij"), -1).reshape(-1, 3)
    pos = ((g[:, None, :] + basis[None]) * a).reshape(-1, 3)
    L = 75.0
    pos = pos - L / 2.0 + 0.4 + np.random.default_rng(9).normal(0, 0.02, pos.shape)
    n = pos.shape[0]
    s = System(pos=pos, mom=np.zeros_like(pos), mass=np.ones(n), charge=np.zeros(n), types=np.zeros(n, np.int32),
               box=(L, L, L), nonbonded=[(0, 0, abi.NB_LJ, (1.0, 1.0, 1.0))])
    rc = 14.5
    assert int(np.floor(L / (rc + 0.3))) == 5 and n <= 40 * 125
    c = cfg(rc, 0.3, engine=engine, coulomb=abi.COUL_CUTOFF, skin=0.3, T=120.272)
    co = cfg(rc, 0.3, coulomb=abi.COUL_CUTOFF, T=120.272)
    o = oracle_mod.Oracle(s, co)
    Fo, Uo = o.compute_forces()
    del o
    e = Engine(s, c)
    Fe, Ue = e.compute_forces()
    if engine == abi.ENGINE_FP64:
        scale = np.maximum(np.linalg.norm(Fo, axis=1), np.sqrt(np.mean(np.sum(Fo ** 2, 1))))
        assert (np.linalg.norm(Fe - Fo, axis=1) / scale).max() < 1e-12
    else:
        assert rms_rel(Fe, Fo) < 1e-5, rms_rel(Fe, Fo)
    assert abs(Ue.astype(np.float64).sum() - Uo.astype(np.float64).sum()) <= 1e-5 * np.abs(Uo).sum()
    th0 = e.thermo()
    e.step(50)
    th1 = e.thermo()
    assert np.isfinite(th1["Etot"])
    assert abs(th1["Etot"] - th0["Etot"]) <= 1e-3 * abs(th0["Etot"]), (th0["Etot"], th1["Etot"])
    e.close()
