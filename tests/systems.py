"""Small seeded systems shared by the parity tests (sizes the oracle finishes
in well under a second)."""
import numpy as np

from moltwister_amd import abi, config, workloads as W


def cfg(rc, shell, ensemble=abi.NVE, dt=0.5, T=298.0, coulomb=abi.COUL_AUTO, engine=abi.ENGINE_EXACT, skin=0.3):
    c = config.defaults()
    c.cutoff_radius = rc
    c.neigh_shell = shell
    c.ensemble = ensemble
    c.timestep = dt
    c.temperature = T
    c.coulomb = coulomb
    c.engine = engine
    c.skin = skin
    return c


# name -> (builder, rc, shell)
SYSTEMS = {
    "fcc_2cells": (lambda: W.fcc_lj(5), 2.5, 0.3),          # 2 cells/axis: duplicated neighbours
    "fcc_4000": (lambda: W.fcc_lj(10), 2.5, 0.3),           # configuration 1 (5^3 cells)
    "binary_lj": (lambda: W.binary_lj(1500, seed=11), 2.5, 0.3),
    "buckingham": (lambda: W.binary_lj(800, seed=12, buckingham=True), 2.5, 0.3),
    "water_300": (lambda: W.water_box(300, seed=5), 6.0, 0.5),
    "chains": (lambda: W.chain_molecules(30, 6, 24.0, seed=4), 8.0, 0.0),
    "chains_rc7": (lambda: W.chain_molecules(30, 6, 24.0, seed=4), 7.0, 0.0),   # >= 3 cells of rc+skin
}


def make(name, **kw):
    b, rc, sh = SYSTEMS[name]
    return b(), cfg(rc, sh, **kw)
