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


def chains_class2():
    """The chain molecules with the other MDFF kinds: Class2 angles, Fourier4t
    dihedrals and an LJBuck pair (whose reference force is zero)."""
    s = W.chain_molecules(30, 6, 24.0, seed=4)
    s.angle_ff = [(abi.ANGLE_CLASS2, (114.0, 250.0, -30.0, 20.0, 5.0, 1.54, 1.54, 8.0, 8.0, 1.54, 1.54))]
    s.dihedral_ff = [(abi.DIH_FOURIER4T, (2.0, -0.5, 3.0, 0.1))]
    s.nonbonded = s.nonbonded[:2] + [(1, 1, abi.NB_LJBUCK, (0, 0.3908, 3.95))]
    return s


# name -> (builder, rc, shell)
SYSTEMS = {
    "fcc_2cells": (lambda: W.fcc_lj(5), 2.5, 0.3),          # 2 cells/axis: duplicated neighbours
    "fcc_4000": (lambda: W.fcc_lj(10), 2.5, 0.3),           # configuration 1 (5^3 cells)
    "binary_lj": (lambda: W.binary_lj(1500, seed=11), 2.5, 0.3),
    "buckingham": (lambda: W.binary_lj(800, seed=12, buckingham=True), 2.5, 0.3),
    "water_300": (lambda: W.water_box(300, seed=5), 6.0, 0.5),
    "chains": (lambda: W.chain_molecules(30, 6, 24.0, seed=4), 8.0, 0.0),
    "chains_rc7": (lambda: W.chain_molecules(30, 6, 24.0, seed=4), 7.0, 0.0),   # >= 3 cells of rc+skin
    "chains_class2": (chains_class2, 7.0, 0.0),
}


def make(name, **kw):
    b, rc, sh = SYSTEMS[name]
    return b(), cfg(rc, sh, **kw)
