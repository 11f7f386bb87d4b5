"""GPU checks of the batch engine (many independent LJ replicas: a persistent
kernel with a workgroup per replica, or, for replicas too large for it, one
captured step graph of many kernels).

Forces of every replica are checked against an O(N^2) numpy evaluation of the
same tabulated force (the table of mtmd_profile, linear interpolation as in
FunctorCalcForce.cpp:457-512) with the minimum image, which also covers boxes
with fewer than three cells per axis (where the reference search visits cells
twice); replicas with three or more cells are also checked against the oracle.
The MIXED arithmetic bar applies: force RMS <= 1e-5 relative.
"""
import numpy as np
import pytest

from moltwister_amd import abi, config, workloads as W
from moltwister_amd.engine import BatchEngine, Engine, profile

pytestmark = pytest.mark.gpu


def lj_cfg(T_star, ensemble=abi.NVT):
    c = config.defaults()
    c.cutoff_radius, c.neigh_shell, c.skin = 2.5, 0.3, 0.3
    c.timestep, c.temperature, c.temperature_relax = 0.5, T_star * 120.272, 20.0
    c.ensemble = ensemble
    c.coulomb = abi.COUL_CUTOFF
    return c


def table_forces(s, rc=2.5, chunk=512, energy=False):
    """Brute-force sum of the tabulated pair force (and with energy=True also
    the total tabulated pair energy) over minimum images, per type pair (the
    system's non-bonded entries), rows in chunks"""
    tabs = {}
    for ta, tb, kind, prm in s.nonbonded:
        x, f, e = profile(0, kind, prm, 0.01, float(np.float32(rc)))
        tabs[(ta, tb)] = tabs[(tb, ta)] = (x.astype(np.float64), f.astype(np.float64), e.astype(np.float64))
    L = np.array(s.box)
    F = np.zeros((s.n_atoms, 3))
    U = 0.0
    for i0 in range(0, s.n_atoms, chunk):
        d = s.pos[i0:i0 + chunk, None, :] - s.pos[None, :, :]
        d -= L * np.round(d / L)
        r = np.sqrt(np.sum(d * d, axis=2))
        rows = np.arange(i0, min(i0 + chunk, s.n_atoms))
        r[rows - i0, rows] = 1e9
        fr = np.zeros_like(r)
        for (ta, tb), (x, f, e) in tabs.items():
            m = (s.types[rows][:, None] == ta) & (s.types[None, :] == tb)
            delta = (x[-1] - x[0]) / 999.0
            idx = np.floor((r - x[0]) / delta).astype(np.int64)
            ok = m & (r < np.float32(rc)) & (idx >= 0) & (idx < 999)
            i = np.clip(idx, 0, 998)
            w = (r - x[i]) / (x[i + 1] - x[i])
            fr += np.where(ok, f[i] + (f[i + 1] - f[i]) * w, 0.0)
            if energy:
                U += 0.5 * float(np.sum(np.where(ok, e[i] + (e[i + 1] - e[i]) * w, 0.0)))
        F[rows] = np.sum(d * (fr / r)[:, :, None], axis=1)
    return (F, U) if energy else F


def rms_rel(F, Fr):
    return float(np.sqrt(np.sum((F - Fr) ** 2) / np.sum(Fr ** 2)))


def small_batch():
    # two small boxes (2 cells per axis: 6.35 and 6.84 sigma) and two with 4-5 cells
    specs = [(256, 1.0, 0.7), (300, 0.9, 1.2), (1500, 0.8, 0.9), (2200, 0.65, 1.4)]
    systems = [W.replica_lj(n, rho, T, seed=21 + i) for i, (n, rho, T) in enumerate(specs)]
    return systems, [t for _, _, t in specs]


def test_batch_forces_vs_brute_force_and_oracle(oracle_mod):
    print("case batch forces", flush=True)
    systems, temps = small_batch()
    b = BatchEngine(systems, [lj_cfg(t) for t in temps])
    for i, s in enumerate(systems):
        F, _ = b.get_forces(i)
        Fr = table_forces(s)
        assert rms_rel(F, Fr) < 1e-5, (i, rms_rel(F, Fr))
        if s.box[0] >= 3 * 2.8:
            o = oracle_mod.Oracle(s, lj_cfg(temps[i]))
            Fo, _ = o.compute_forces()
            assert rms_rel(F, Fo) < 1e-5, (i, rms_rel(F, Fo))
    b.close()


def test_batch_graph_equals_separate_launches():
    """The whole-batch launch and per-replica launches run the same code on
    the same data: identical states."""
    print("case batch graph vs separate", flush=True)
    systems, temps = small_batch()
    cfgs = [lj_cfg(t) for t in temps]
    a, b = BatchEngine(systems, cfgs), BatchEngine(systems, cfgs)
    a.step(40)
    b.step_separate(40)
    for i in range(len(systems)):
        pa, ma = a.get_state(i)
        pb, mb = b.get_state(i)
        assert np.array_equal(pa, pb) and np.array_equal(ma, mb), i
    ta, tb = a.thermo(), b.thermo()
    for x, y in zip(ta, tb):
        assert x["Etot"] == y["Etot"] and x["T"] == y["T"]
    a.close()
    b.close()


def test_batch_nvt_replica_matches_single_system_engine():
    """A replica of the batch follows the same NVT trajectory as the MIXED
    single-system engine on the same system (same chain at rest, same
    momenta), up to FP32 pair rounding, over 200 steps."""
    print("case batch nvt vs single", flush=True)
    systems, temps = small_batch()
    i = 2
    b = BatchEngine(systems, [lj_cfg(t) for t in temps])
    c = lj_cfg(temps[i])
    c.engine = abi.ENGINE_MIXED
    e = Engine(systems[i], c)
    for _ in range(4):
        b.step(50)
        e.step(50)
        tb, te = b.thermo()[i], e.thermo()
        assert abs(tb["T"] - te["T"]) < 2e-3 * te["T"], (tb["T"], te["T"])
        assert abs(tb["Etot"] - te["Etot"]) < 2e-3 * abs(te["Etot"]), (tb["Etot"], te["Etot"])
    b.close()


def test_batch_nvt_holds_each_replica_temperature():
    """Each replica's mean temperature over 1000 steps (after 600 steps of
    equilibration; the chain's relaxation time is 40 steps) at its own target:
    within four standard errors of the trace's mean (a sample every 10 steps,
    one independent sample per 100 steps) plus 1% for the time step's bias."""
    print("case batch temperatures", flush=True)
    systems, temps = W.replicas(12, seed=3, n_min=256, n_max=1200)
    b = BatchEngine(systems, [lj_cfg(t) for t in temps])
    b.step(600)
    tr = []
    for _ in range(100):
        b.step(10)
        tr.append([th["T"] / 120.272 for th in b.thermo()])
    tr = np.array(tr)
    assert np.all(np.isfinite(tr))
    for i, t in enumerate(temps):
        m, se = tr[:, i].mean(), tr[:, i].std() / np.sqrt(len(tr) / 10)
        assert abs(m - t) < 4 * se + 0.01 * t, (i, systems[i].n_atoms, m, t, se)
    b.close()


def test_batch_autotune_and_nve_energy():
    """After the autotuner, an NVE batch started from thermalised states keeps
    every replica's total energy (truncation jumps of the unshifted cutoff
    allowed for)."""
    print("case batch autotune + nve", flush=True)
    systems, temps = small_batch()
    warm = BatchEngine(systems, [lj_cfg(t) for t in temps])
    warm.step(400)
    for i, s in enumerate(systems):
        s.pos, s.mom = warm.get_state(i)
    warm.close()
    b = BatchEngine(systems, [lj_cfg(t, abi.NVE) for t in temps])
    info = b.autotune(2)
    # every replica fits the persistent kernel: cell shape x workgroup size x lanes per atom
    assert info["persistent"] and info["block"] in (256, 512, 1024) and info["lanes"] in (1, 2, 4), info
    assert info["cell_split"] in (1, 2), info
    e0 = [t["Etot"] for t in b.thermo()]
    b.step(200)
    e1 = [t["Etot"] for t in b.thermo()]
    for x, y in zip(e0, e1):
        assert abs(y - x) < 5e-3 * abs(x), (x, y)
    b.close()


def test_batch_kernel_path_agrees(monkeypatch):
    """The graph path of many kernels (forced by MTMD_BATCH_KERNELS) and the
    persistent kernel: the same forces up to FP32 summation order, the same
    NVT trajectory up to its growth over 100 steps."""
    print("case batch persistent vs kernels", flush=True)
    systems, temps = small_batch()
    cfgs = [lj_cfg(t) for t in temps]
    a = BatchEngine(systems, cfgs)
    monkeypatch.setenv("MTMD_BATCH_KERNELS", "1")
    k = BatchEngine(systems, cfgs)
    monkeypatch.delenv("MTMD_BATCH_KERNELS")
    assert k.info()["list_block"] in (64, 128, 256)
    for i in range(len(systems)):
        Fa, Ua = a.get_forces(i)
        Fk, Uk = k.get_forces(i)
        assert rms_rel(Fa, Fk) < 1e-6, (i, rms_rel(Fa, Fk))
        assert abs(float(np.sum(Ua)) - float(np.sum(Uk))) < 1e-5 * abs(float(np.sum(Uk))), i
    a.step(100)
    k.step(100)
    for x, y in zip(a.thermo(), k.thermo()):
        assert abs(x["T"] - y["T"]) < 2e-3 * y["T"] and abs(x["Etot"] - y["Etot"]) < 2e-3 * abs(y["Etot"]), (x, y)
    a.close()
    k.close()


def test_batch_two_cells_per_radius(monkeypatch):
    """The persistent kernel with y and z cells of half a list radius (a
    five-row stencil on those axes) for the replicas large enough: forces
    against the brute-force table sum, and the same NVE run as with whole
    cells up to FP32 summation order."""
    print("case batch cell split 2", flush=True)
    systems, temps = small_batch()
    cfgs = [lj_cfg(t, abi.NVE) for t in temps]
    a = BatchEngine(systems, cfgs)
    monkeypatch.setenv("MTMD_BATCH_CELL_SPLIT", "2")
    b = BatchEngine(systems, cfgs)
    monkeypatch.delenv("MTMD_BATCH_CELL_SPLIT")
    assert b.tuning()["cell_split"] == 2 and a.tuning()["cell_split"] == 1
    for i, s in enumerate(systems):
        F, _ = b.get_forces(i)
        Fa, _ = a.get_forces(i)
        assert rms_rel(F, table_forces(s)) < 1e-5, i
        assert rms_rel(F, Fa) < 1e-6, i
    a.step(100)
    b.step(100)
    for x, y in zip(a.thermo(), b.thermo()):
        assert abs(x["Etot"] - y["Etot"]) < 1e-4 * abs(x["Etot"]) and abs(x["T"] - y["T"]) < 1e-3 * x["T"], (x, y)
    a.close()
    b.close()


def test_batch_list_regrow(monkeypatch, capfd):
    """List growth inside a call (MDFFMatrices.cpp:163-178 grows and builds
    again): with MTMD_BATCH_LIST_SLACK=0 the rows start at 16 words and any
    row longer than the longest one so far grows them, so builds inside a call
    overflow as the rows fluctuate.
    The first build grows until it fits and then equals a batch with room to
    spare, bit for bit; calls that overflow restore their starting state and
    run again, ending on the same trajectory up to the summation order of
    lists built at other steps."""
    print("case batch list regrow", flush=True)
    systems, temps = small_batch()
    cfgs = [lj_cfg(t) for t in temps]
    warm = BatchEngine(systems, cfgs)     # from a liquid: the lattice start has the longest rows
    warm.step(400)
    for i, s in enumerate(systems):
        s.pos, s.mom = warm.get_state(i)
    warm.close()
    a = BatchEngine(systems, cfgs)
    monkeypatch.setenv("MTMD_BATCH_LIST_SLACK", "0")
    monkeypatch.setenv("MTMD_BATCH_LOG", "1")
    b = BatchEngine(systems, cfgs)
    for i in range(len(systems)):
        assert np.array_equal(a.get_forces(i)[0], b.get_forces(i)[0]), i
    grown = ""
    for _ in range(40):                   # until a call has overflowed (rows fluctuate as atoms move)
        a.step(50)
        b.step(50)
        grown = capfd.readouterr().err
        if "lists grown" in grown:
            break
    assert "lists grown" in grown
    b.step_separate(5)
    a.step_separate(5)
    for x, y in zip(a.thermo(), b.thermo()):
        assert abs(x["T"] - y["T"]) < 1e-4 * x["T"] and abs(x["Etot"] - y["Etot"]) < 1e-4 * abs(x["Etot"]), (x, y)
    a.close()
    b.close()


def binary_replica(n, rho, seed):
    """a two-type LJ replica (80:20, eps/sigma AA 1/1, AB 1.5/0.8, BB 0.5/0.88),
    type 0 always A, so that every replica shares the force-field list"""
    s = W.replica_lj(n, rho, 1.0, seed)
    rng = np.random.default_rng(seed + 1)
    s.types = (rng.random(n) < 0.2).astype(np.int32)
    s.type_names = ["A", "B"]
    s.nonbonded = [(0, 0, abi.NB_LJ, (1.0, 1.0, 1.0)), (0, 1, abi.NB_LJ, (1.5, 0.8, 1.0)),
                   (1, 1, abi.NB_LJ, (0.5, 0.88, 1.0))]
    return s


def test_batch_two_types_and_largest_replica():
    """The persistent kernel at its limits: two atom types (a table of four
    type-pair rows in LDS), a replica of 4096 atoms (the staging's capacity),
    and a box under three cells per axis (minimum image per pair), against the
    brute-force table sum per type pair; then 200 NVE steps keep each
    replica's energy."""
    print("case batch two types, 4096 atoms", flush=True)
    systems = [binary_replica(4096, 1.0, 31), binary_replica(300, 0.9, 32), binary_replica(2000, 0.7, 33)]
    warm = BatchEngine(systems, [lj_cfg(1.0) for _ in systems])
    assert warm.tuning()["persistent"]
    for i, s in enumerate(systems):
        F, Ua = warm.get_forces(i)
        Fr, Ur = table_forces(s, energy=True)
        assert rms_rel(F, Fr) < 1e-5, (i, rms_rel(F, Fr))
        assert abs(float(np.sum(Ua)) - Ur) < 1e-5 * abs(Ur), (i, float(np.sum(Ua)), Ur)
    # from a liquid (the lattice start melts, and pairs crossing the unshifted
    # cutoff move Etot): 200 NVE steps keep each replica's energy
    warm.step(400)
    for i, s in enumerate(systems):
        s.pos, s.mom = warm.get_state(i)
    warm.close()
    b = BatchEngine(systems, [lj_cfg(1.0, abi.NVE) for _ in systems])
    e0 = [t["Etot"] for t in b.thermo()]
    b.step(200)
    for x, t in zip(e0, b.thermo()):
        assert abs(t["Etot"] - x) < 5e-3 * abs(x), (x, t["Etot"])
    b.close()


def test_batch_rejects_charges():
    systems, temps = small_batch()
    systems[1].charge[:] = 0.1
    with pytest.raises(Exception, match="charges"):
        BatchEngine(systems, [lj_cfg(t) for t in temps])
