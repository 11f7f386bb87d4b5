"""NPT (the MTK barostat of CVelVerlet with the NH_P chain, VelVerlet.cpp:57-71,
124-174,192-211; MDLoop.cpp:63-83,124-138) on the HIP engines: EXACT against
the oracle step by step, FP64 and MIXED against EXACT over trajectories with
list rebuilds while the box moves, and mtmd_run's NPT table."""
import numpy as np
import pytest

from moltwister_amd import abi
from moltwister_amd.engine import Engine, run

from systems import make

pytestmark = pytest.mark.gpu

CONV_T, CONV_t = 120.272, 100.0


def npt_setup(name, engine=abi.ENGINE_EXACT, pressure=2000.0):
    s, c = make(name, ensemble=abi.NPT, engine=engine)
    c.temperature = 86.596 if name.startswith("fcc") or name == "binary_lj" else 298.0
    c.pressure = pressure
    c.pressure_relax = 5000.0          # W = T tau_P^2 (no factor N): a light barostat needs a long tau_P
    c.temperature_relax = 50.0
    c.cutoff_force = 0.0
    T = c.temperature / CONV_T
    tau, tauP = c.temperature_relax / CONV_t, c.pressure_relax / CONV_t
    Q = np.array([3 * s.n_atoms, 1, 1, 1]) * T * tau * tau
    W = T * tauP * tauP
    chain_t = (np.array([0.1, 0.2, 0.3, 0.4]), np.array([0.3, -0.2, 0.1, 0.05]) * Q / tau)
    chain_p = (0.0, 0.4 * W / tauP, np.array([0.05, 0.1, -0.1, 0.2]),
               np.array([-0.2, 0.1, 0.05, 0.02]) * T * tauP)
    return s, c, chain_t, chain_p


def start(x, chain_t, chain_p):
    x.set_thermostat(*chain_t)
    x.set_barostat(*chain_p)


def _rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


@pytest.mark.parametrize("name", ["fcc_4000", "water_300"])
def test_exact_npt_steps_match_oracle(oracle_mod, name):
    print(f"case npt exact {name}", flush=True)
    s, c, ct, cp = npt_setup(name)
    o, e = oracle_mod.Oracle(s, c), Engine(s, c)
    start(o, ct, cp)
    start(e, ct, cp)
    o.step(1)
    e.step(1)
    po, mo = o.get_state()
    pg, mg = e.get_state()
    assert np.allclose(pg, po, rtol=0, atol=1e-12 * max(1.0, np.abs(po).max()))
    assert np.allclose(mg, mo, rtol=1e-11, atol=1e-12 * np.abs(mo).max())
    bo, bg = o.get_barostat(), e.get_barostat()
    assert _rel(bg["eps"], bo["eps"]) < 1e-12 and _rel(bg["p_eps"], bo["p_eps"]) < 1e-10
    assert np.allclose(bg["box"], bo["box"], rtol=1e-14, atol=0)
    assert np.allclose(bg["p_eta"], bo["p_eta"], rtol=1e-10)
    # a short trace: T, P, V, U pointwise
    for _ in range(5):
        o.step(10)
        e.step(10)
        to, tg = o.thermo(), e.thermo()
        for k in ("T", "P", "V", "U", "K", "Etot"):
            assert _rel(tg[k], to[k]) < 1e-8, (k, tg[k], to[k])
    bo, bg = o.get_barostat(), e.get_barostat()
    assert abs(bo["eps"]) > 1e-5                     # the box did move
    assert _rel(bg["eps"], bo["eps"]) < 1e-8


@pytest.mark.parametrize("engine", [abi.ENGINE_FP64, abi.ENGINE_MIXED])
def test_sorted_engines_npt_follow_exact(engine):
    print(f"case npt fast {engine}", flush=True)
    s, c, ct, cp = npt_setup("binary_lj", pressure=5.0e4)
    s2, c2, _, _ = npt_setup("binary_lj", engine=engine, pressure=5.0e4)
    ex, fa = Engine(s, c), Engine(s2, c2)
    start(ex, ct, cp)
    start(fa, ct, cp)
    tol = 1e-9 if engine == abi.ENGINE_FP64 else 2e-5
    for _ in range(4):
        ex.step(25)
        fa.step(25)
        te, tf = ex.thermo(), fa.thermo()
        for k in ("T", "P", "V", "Etot"):
            assert _rel(tf[k], te[k]) < tol * 50, (k, tf[k], te[k])
        be, bf = ex.get_barostat(), fa.get_barostat()
        assert np.allclose(bf["box"], be["box"], rtol=tol, atol=0)
    pe, _ = ex.get_state()
    pf, _ = fa.get_state()
    L = be["box"]
    d = pf - pe
    d -= L * np.rint(d / L)
    assert np.abs(d).max() < 1e3 * tol
    assert abs(be["eps"]) > 1e-4                     # the box changed by more than the skin allows unbuilt
    assert fa.stats()["rebuilds"] > 1


def test_npt_volume_wrap_and_thermo():
    s, c, ct, cp = npt_setup("fcc_4000", engine=abi.ENGINE_MIXED, pressure=-2000.0)
    e = Engine(s, c)
    start(e, ct, cp)
    e.step(60)
    b = e.get_barostat()
    th = e.thermo()
    V0 = np.prod(s.box)
    assert th["V"] == pytest.approx(V0 * np.exp(3.0 * b["eps"]), rel=1e-14)
    assert np.allclose(b["box"], np.array(s.box) * np.exp(b["eps"]), rtol=1e-12)
    pos, _ = e.get_state()
    assert np.all(pos >= -b["box"] / 2) and np.all(pos < b["box"] / 2)


def test_run_npt_table(tmp_path):
    s, c, _, _ = npt_setup("fcc_4000", engine=abi.ENGINE_FP64)
    c.n_steps, c.out_stride = 201, 100
    c.info_file = str(tmp_path / "npt.txt").encode()
    c.include_dcd = 0
    run(s, c)
    lines = open(tmp_path / "npt.txt", "rb").read().decode().split("\r\n")
    assert any("Ensemble = NPT" in ln for ln in lines)
    assert any(ln.startswith("\t Pressure, P = 2000 atm") for ln in lines)
    assert any(ln.startswith("\t NH P relax, tau = 5000 fs") for ln in lines)
    rows = [ln.split() for ln in lines if ln.startswith("\t") and ln.strip()[:1].isdigit()]
    assert [int(r[0]) for r in rows] == [0, 100, 200]
    V = [float(r[4]) for r in rows]
    assert len(set(V)) == 3                          # the volume column moves with the barostat
