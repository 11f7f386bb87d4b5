"""The FP64/MIXED step driver launches a step in several equivalent ways:
every kernel directly, or as a replayed HIP graph of a plain step (MTMD_GRAPHS);
and, for MIXED NVE on the tile path, with the half kicks and drift of the
velocity Verlet step (VelVerlet.cpp:72-122) inside the pair-force pass or in a
kernel of their own (MTMD_FUSE). All of them run the same per-atom arithmetic
in the same order, so the trajectories agree bit for bit: positions, momenta,
forces, the list rebuild count and the thermodynamic sums, over calls of
uneven length that take in the spatial re-sort every 100 steps (the graph
state changes there) and list rebuilds."""
import os

import numpy as np
import pytest

from moltwister_amd import abi, config
from moltwister_amd import workloads as W
from moltwister_amd.engine import Engine

pytestmark = pytest.mark.gpu

N = 131072
CALLS = (1, 57, 140, 2, 60)          # 260 steps: re-sorts at steps 100 and 200


def cfg(engine, ensemble):
    c = config.defaults()
    c.engine = engine
    c.ensemble = ensemble
    c.cutoff_radius = 2.5
    c.neigh_shell = 0.3
    c.skin = 0.3
    c.timestep = 0.5
    c.temperature = 120.272
    c.cutoff_force = 1000.0
    c.coulomb = abi.COUL_CUTOFF
    return c


@pytest.fixture(scope="module")
def liquid():
    s = W.binary_lj_large(N, tiles=2, seed=5)
    e = Engine(s, cfg(abi.ENGINE_MIXED, abi.NVE))
    # a liquid near the thermostat's temperature: the Nose-Hoover chain of the
    # reference diverges on the heating of a barely relaxed start
    e.optimize(100, 2e-5, 0.0)
    for _ in range(3):
        e.step(100)
        pos, mom = e.get_state()
        e.set_state(mom=mom * np.sqrt(120.272 / e.thermo()["T"]))
    s.pos, s.mom = e.get_state()
    e.close()
    return s


def trajectory(s, c, graphs, fuse):
    old = {k: os.environ.get(k) for k in ("MTMD_GRAPHS", "MTMD_FUSE")}
    os.environ["MTMD_GRAPHS"] = "1" if graphs else "0"
    os.environ["MTMD_FUSE"] = "1" if fuse else "0"
    try:
        e = Engine(s, c)
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    out = []
    for n in CALLS:
        e.step(n)
        pos, mom = e.get_state()
        F, _ = e.get_forces()
        th = e.thermo()
        out.append((pos, mom, F, th["T"], th["U"], e.stats()["rebuilds"]))
    e.close()
    return out


def same(a, b):
    for k, (x, y) in enumerate(zip(a, b)):
        for name, u, v in zip(("pos", "mom", "F"), x[:3], y[:3]):
            assert np.array_equal(u, v), (k, name, np.abs(u - v).max())
        assert x[3:] == y[3:], (k, x[3:], y[3:])


@pytest.mark.parametrize("engine,ensemble", [(abi.ENGINE_MIXED, abi.NVE), (abi.ENGINE_MIXED, abi.NVT),
                                             (abi.ENGINE_FP64, abi.NVE)])
def test_step_paths_agree(liquid, engine, ensemble):
    print(f"case step paths engine={engine} ensemble={ensemble}", flush=True)
    c = cfg(engine, ensemble)
    ref = trajectory(liquid, c, graphs=False, fuse=False)
    assert ref[-1][5] > 2                         # rebuilds happened in the run
    variants = [(True, False)]
    if engine == abi.ENGINE_MIXED and ensemble == abi.NVE:
        variants += [(False, True), (True, True)]
    for graphs, fuse in variants:
        print(f"  graphs={graphs} fuse={fuse}", flush=True)
        same(trajectory(liquid, c, graphs, fuse), ref)
