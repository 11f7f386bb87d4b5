"""Two parts of the integrator path against the oracle with the part acting:

- the force cap (CVelVerlet::cutForces, VelVerlet.cpp:87,96-122): every force
  component beyond +-Fcut is set to +-Fcut before the second half kick. A pair
  placed at 0.8 sigma and Fcut = 50 make it act on the first step and on the
  steps after. EXACT is bit-identical to the oracle (state and capped forces);
  FP64 per atom within 1e-12 of max(|F_i|, F_rms); MIXED capped forces within
  the force-RMS bar 1e-5;
- the descent of `moldyn optimizeenergy` (CMDLoop::runOptimization,
  MDLoop.cpp:143-173, with CSimulationBox::gradientDescentStep,
  SimulationBox.cpp:197-208): r += lr F for at most max_steps steps, stopping
  when |U(n+1) - U(n)| < accuracy after step out_stride. EXACT gives the
  oracle's positions bit for bit and stops at the same step."""
import numpy as np
import pytest

from moltwister_amd import abi
from moltwister_amd.engine import Engine

from systems import make

pytestmark = pytest.mark.gpu

FCUT = 50.0


def capped_system(engine=abi.ENGINE_EXACT):
    s, c = make("binary_lj", engine=engine, coulomb=abi.COUL_CUTOFF)
    c.cutoff_force = FCUT
    # atom 1 at 0.8 sigma from atom 0 along x: an overlapping pair
    pos = s.pos.copy()
    pos[1] = pos[0] + np.array([0.8, 0.0, 0.0])
    s.pos = pos
    return s, c


def per_atom_rel(F, Fr):
    scale = np.maximum(np.linalg.norm(Fr, axis=1), np.sqrt(np.mean(np.sum(Fr ** 2, 1))))
    return np.linalg.norm(F - Fr, axis=1) / scale


def test_force_cap_acts_exact(oracle_mod):
    print("case force cap exact", flush=True)
    s, c = capped_system()
    o, e = oracle_mod.Oracle(s, c), Engine(s, c)
    F0, _ = o.compute_forces()                       # uncapped, initial positions
    assert np.abs(F0).max() > 10 * FCUT              # the overlapping pair
    assert np.count_nonzero(np.abs(F0) > FCUT) > 2
    for nsteps in (1, 4, 15):
        o.step(nsteps)
        e.step(nsteps)
        po, mo = o.get_state()
        pe, me = e.get_state()
        assert np.array_equal(po, pe), (nsteps, np.abs(po - pe).max())
        assert np.array_equal(mo, me), (nsteps, np.abs(mo - me).max())
        Fo, _ = o.get_forces()
        Fe, _ = e.get_forces()
        assert np.array_equal(Fo, Fe), (nsteps, np.abs(Fo - Fe).max())
        assert np.abs(Fe).max() <= FCUT
    # the cap acted on the forces the integrator used
    assert np.count_nonzero(np.abs(Fe) == FCUT) >= 1


@pytest.mark.parametrize("engine", [abi.ENGINE_FP64, abi.ENGINE_MIXED])
def test_force_cap_acts_fast(oracle_mod, engine):
    print(f"case force cap engine={engine}", flush=True)
    s, c = capped_system()
    sf, cf = capped_system(engine)
    o, e = oracle_mod.Oracle(s, c), Engine(sf, cf)
    o.step(1)
    e.step(1)
    Fo, _ = o.get_forces()
    Fe, _ = e.get_forces()
    capped = np.abs(Fo) == FCUT
    assert capped.sum() >= 1
    assert np.abs(Fe).max() <= FCUT
    po, mo = o.get_state()
    pe, me = e.get_state()
    if engine == abi.ENGINE_FP64:
        assert per_atom_rel(Fe, Fo).max() < 1e-12
        assert np.array_equal(Fe[capped], Fo[capped])
        assert np.abs(pe - po).max() < 1e-12
        assert np.abs(me - mo).max() < 1e-11 * np.abs(mo).max()
    else:
        rms = float(np.sqrt(np.sum((Fe - Fo) ** 2) / np.sum(Fo ** 2)))
        assert rms < 1e-5, rms
        # a component whose uncapped oracle force is well beyond the threshold
        # is capped here too, to the same value
        Fu, _ = o.compute_forces()
        deep = np.abs(Fu) > 1.01 * FCUT
        assert deep.sum() >= 1
        assert np.array_equal(Fe[deep], Fo[deep])
        assert np.abs(pe - po).max() < 1e-6


@pytest.mark.parametrize("name,lr,accuracy,stop", [("binary_lj", 2e-5, 100.0, 38), ("water_300", 1e-4, 100.0, 34),
                                                   ("binary_lj", 2e-5, 0.0, 60)])
def test_descent_matches_oracle(oracle_mod, name, lr, accuracy, stop):
    print(f"case descent {name} accuracy={accuracy}", flush=True)
    s, c = make(name)
    c.out_stride = 10
    o, e = oracle_mod.Oracle(s, c), Engine(s, c)
    n_o = o.optimize(60, lr, accuracy, 10)
    n_e = e.optimize(60, lr, accuracy)
    assert n_o == stop
    assert n_e == n_o
    po, _ = o.get_state()
    pe, _ = e.get_state()
    assert np.array_equal(po, pe), np.abs(po - pe).max()
    assert np.abs(pe - s.pos).max() > 1e-3
