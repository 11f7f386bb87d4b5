"""GPU parity of the EXACT engine (reference algorithm) against the oracle.

Bar: cell assignments, cell lists, neighbour lists and per-atom forces and
energies bit-identical to the oracle (which is bit-identical to the
reference's own CPU sources, tests/test_oracle_ref.py). Thermostat and thermo
reductions use a different (parallel) summation order, so the traces are
compared at 1e-12 relative for a single step and 1e-8 over 100 steps.
"""
import numpy as np
import pytest

from moltwister_amd import abi
from moltwister_amd.engine import Engine

from systems import SYSTEMS, make

pytestmark = pytest.mark.gpu

NAMES = list(SYSTEMS)


@pytest.fixture(scope="module")
def P(oracle_mod):
    return oracle_mod


@pytest.mark.parametrize("name", NAMES)
def test_cells_bit_exact(P, name):
    print(f"case cells {name}", flush=True)
    s, c = make(name)
    o, e = P.Oracle(s, c), Engine(s, c)
    co, ce = o.build_cells(), e.build_cells()
    for key in ("nc", "ixyz", "first", "count", "list"):
        assert np.array_equal(co[key], ce[key]), key


@pytest.mark.parametrize("name", NAMES)
def test_neighbors_bit_exact(P, name):
    print(f"case neighbours {name}", flush=True)
    s, c = make(name)
    o, e = P.Oracle(s, c), Engine(s, c)
    no, lo = o.build_neighbors()
    ne, le = e.build_neighbors()
    assert np.array_equal(no, ne)
    for k in range(s.n_atoms):
        assert np.array_equal(lo[k, :no[k]], le[k, :ne[k]]), k


@pytest.mark.parametrize("name", NAMES)
def test_forces_bit_exact(P, name):
    print(f"case forces {name}", flush=True)
    s, c = make(name)
    o, e = P.Oracle(s, c), Engine(s, c)
    Fo, Uo = o.compute_forces()
    Fe, Ue = e.compute_forces()
    assert np.array_equal(Fo, Fe), np.abs(Fo - Fe).max()
    assert np.array_equal(Uo, Ue), np.abs(Uo - Ue).max()


@pytest.mark.parametrize("name", ["binary_lj", "water_300"])
def test_forces_after_random_displacement(P, name):
    """Positions moved off the generator's grid and partly outside the box
    (clamped cells, minimum image): still bit-identical."""
    print(f"case displaced forces {name}", flush=True)
    s, c = make(name)
    rng = np.random.default_rng(7)
    pos = s.pos + rng.normal(0, 0.05, s.pos.shape)
    pos[:5] += np.array(s.box) * 0.51      # a few atoms beyond +L/2
    o, e = P.Oracle(s, c), Engine(s, c)
    o.set_state(pos=pos)
    e.set_state(pos=pos)
    Fo, Uo = o.compute_forces()
    Fe, Ue = e.compute_forces()
    assert np.array_equal(Fo, Fe)
    assert np.array_equal(Uo, Ue)


def _rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


@pytest.mark.parametrize("name,ensemble", [("fcc_4000", abi.NVE), ("binary_lj", abi.NVT), ("water_300", abi.NVT)])
def test_one_step_state(P, name, ensemble):
    print(f"case one step {name} {ensemble}", flush=True)
    s, c = make(name, ensemble=ensemble)
    o, e = P.Oracle(s, c), Engine(s, c)
    if ensemble == abi.NVT:
        eta = np.array([0.1, 0.2, 0.3, 0.4])
        Q = np.array([3 * s.n_atoms, 1, 1, 1]) * (c.temperature / 120.272) * (c.temperature_relax / 100.0) ** 2
        pe = np.array([0.3, -0.2, 0.1, 0.05]) * Q / (c.temperature_relax / 100.0)
        o.set_thermostat(eta, pe)
        e.set_thermostat(eta, pe)
    o.step(1)
    e.step(1)
    po, mo = o.get_state()
    pg, mg = e.get_state()
    assert np.allclose(pg, po, rtol=0, atol=1e-12 * max(1.0, np.abs(po).max()))
    assert np.allclose(mg, mo, rtol=1e-11, atol=1e-12 * np.abs(mo).max())
    Fo, _ = o.get_forces()
    Fe, _ = e.get_forces()
    assert np.allclose(Fe, Fo, rtol=1e-9, atol=1e-10 * np.abs(Fo).max())
    to, tg = o.thermo(), e.thermo()
    for k in ("T", "P", "U", "K", "Etot"):
        assert _rel(tg[k], to[k]) < 1e-10, (k, tg[k], to[k])


def test_trace_from_reference_draws(P):
    """A run started as a first `moldyn run` starts: momenta and thermostat
    chain from the reference's own draws (mtmd_reference_draws, seed 1),
    EXACT engine against the oracle over 50 NVT steps, pointwise within 1e-8."""
    print("case trace from reference draws water_300", flush=True)
    from moltwister_amd.engine import reference_draws
    s, c = make("water_300", ensemble=abi.NVT)
    d = reference_draws(s, c, 1)
    s.mom = d["mom"]
    o, e = P.Oracle(s, c), Engine(s, c)
    o.set_thermostat(d["eta_T"], d["p_eta_T"])
    e.set_thermostat(d["eta_T"], d["p_eta_T"])
    for blk in range(5):
        o.step(10)
        e.step(10)
        to, tg = o.thermo(), e.thermo()
        for k in ("T", "P", "U", "K", "Etot"):
            assert _rel(tg[k], to[k]) < 1e-8, (blk, k, tg[k], to[k])


@pytest.mark.parametrize("name,ensemble", [("fcc_4000", abi.NVE), ("water_300", abi.NVT)])
def test_trace_100_steps(P, name, ensemble):
    """Thermo trace over 100 steps, pointwise within 1e-8 relative."""
    print(f"case trace {name} {ensemble}", flush=True)
    s, c = make(name, ensemble=ensemble)
    o, e = P.Oracle(s, c), Engine(s, c)
    if ensemble == abi.NVT:
        eta, pe = e.get_thermostat()
        o.set_thermostat(eta, pe)
    for blk in range(10):
        o.step(10)
        e.step(10)
        to, tg = o.thermo(), e.thermo()
        for k in ("T", "P", "U", "K", "Etot"):
            assert _rel(tg[k], to[k]) < 1e-8, (blk, k, tg[k], to[k])
