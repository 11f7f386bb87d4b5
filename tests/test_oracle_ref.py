"""The oracle against MolTwister's own CPU sources (oracle/_ref/libmtref.so).

The reference build exists only where the MolTwister tree is mounted (it is built
from those sources by oracle/Makefile); elsewhere these tests skip and the
committed fixtures (test_oracle_golden.py) carry the same pins.
Bar: bit-identical (np.array_equal) profiles, cells, neighbour lists,
per-atom forces and energies, NH-chain propagation and CMt helpers.
"""
import numpy as np
import pytest

from moltwister_amd import abi

from systems import SYSTEMS, make

pytestmark = pytest.mark.skipif(
    not __import__("os").path.exists(__import__("oracle.pyoracle", fromlist=["x"]).REF_LIB),
    reason="reference build oracle/_ref/libmtref.so not present (needs the MolTwister source tree)")

PROFILES = [
    (0, abi.NB_LJ, (1.0, 1.0, 1.0), 0.01, 2.5),
    (0, abi.NB_LJ, (0.650194, 3.1655, 1.0), 0.01, 10.0),
    (0, abi.NB_LJ, (1.5, 0.8, 0.5), 0.01, 2.5),
    (0, abi.NB_BUCK, (1000.0, 0.1, 1.0), 0.01, 2.5),
    (0, abi.NB_LJ1208, (0.7, 3.2, 1.0), 0.01, 10.0),
    (1, abi.BOND_HARM, (2215.76, 1.012), 0.01, 10.0),
    (1, abi.BOND_MORSE, (400.0, 2.0, 1.1), 0.01, 10.0),
    (2, abi.ANGLE_HARM, (158.78, 113.24), 0.0, float(np.float32(2 * np.pi))),
    (3, abi.DIH_HARM, (5.0, 1, 3), 0.0, float(np.float32(2 * np.pi))),
    (3, abi.DIH_HARM, (2.0, -1, 2), 0.0, float(np.float32(2 * np.pi))),
    (0, abi.NB_LJBUCK, (0, 1.2, 3.1), 0.01, 10.0),
    (0, abi.NB_LJBUCK, (1, 1000.0, 0.2, 5.0), 0.01, 8.0),
    (2, abi.ANGLE_CLASS2, (112.0, 40.0, -20.0, 10.0, 5.0, 1.1, 1.3, 3.0, 4.0, 1.2, 0.9), 0.0,
     float(np.float32(2 * np.pi))),
    (3, abi.DIH_FOURIER4T, (1.5, -2.0, 0.7, 0.3), 0.0, float(np.float32(2 * np.pi))),
]


@pytest.mark.parametrize("case", PROFILES, ids=lambda c: f"{c[0]}-{c[1]}-{c[2][0]}")
def test_profiles(oracle_mod, case):
    which, kind, prm, x0, x1 = case
    a = oracle_mod.profile(which, kind, prm, x0, x1)
    b = oracle_mod.profile(which, kind, prm, x0, x1, ref=True)
    for u, v in zip(a, b):
        assert np.array_equal(u, v, equal_nan=True)


@pytest.mark.parametrize("name", list(SYSTEMS))
def test_cells_neighbours_forces(oracle_mod, name):
    s, c = make(name)
    o, r = oracle_mod.Oracle(s, c), oracle_mod.Reference(s, c)
    co, cr = o.build_cells(), r.build_cells()
    for key in ("nc", "ixyz", "first", "count", "list"):
        assert np.array_equal(co[key], cr[key]), key
    assert co["err"] == cr["err"]
    no, lo = o.build_neighbors()
    nr, lr = r.build_neighbors()
    assert np.array_equal(no, nr)
    for k in range(s.n_atoms):
        assert np.array_equal(lo[k, :no[k]], lr[k, :nr[k]])
    Fo, Uo = o.compute_forces()
    Fr, Ur = r.compute_forces()
    assert np.array_equal(Fo, Fr)
    assert np.array_equal(Uo, Ur)
    mo, mr = o.mol(), r.mol()
    # molecule labels may differ; the partition must not
    assert len(set(zip(mo, mr))) == len(set(mo)) == len(set(mr))


def test_forces_displaced_and_outside_box(oracle_mod):
    s, c = make("water_300")
    rng = np.random.default_rng(1)
    pos = s.pos + rng.normal(0, 0.1, s.pos.shape)
    pos[:9] += np.array(s.box) * 0.52
    o, r = oracle_mod.Oracle(s, c), oracle_mod.Reference(s, c)
    o.set_state(pos=pos)
    r.set_state(pos=pos)
    Fo, Uo = o.compute_forces()
    Fr, Ur = r.compute_forces()
    assert np.array_equal(Fo, Fr) and np.array_equal(Uo, Ur)


@pytest.mark.parametrize("p_eta0", [0.7, 16.0, -3.0])
def test_nh_chain_propagator(oracle_mod, p_eta0):
    """One CNHChain::propagator call with CFctT, including the p_eta[0] > 15
    reset and the CMt::exp linear branch."""
    s, c = make("water_300", ensemble=abi.NVT)
    o, r = oracle_mod.Oracle(s, c), oracle_mod.Reference(s, c)
    eta = np.array([0.1, 0.2, 0.3, 0.4])
    pe = np.array([p_eta0, -3.0, 2.0, 0.5])
    o.set_thermostat(eta, pe)
    r.set_thermostat(eta, pe)
    o.nh_propagate()
    mom_r = r.nh_propagate()
    _, mom_o = o.get_state()
    assert np.array_equal(mom_o, mom_r)
    for u, v in zip(o.get_thermostat(), r.get_thermostat()):
        assert np.array_equal(u, v)


def test_cmt_helpers(oracle_mod):
    for x in np.concatenate([np.linspace(-5, 5, 401), [1.0, np.nextafter(1.0, 2.0), 0.0, -0.0]]):
        assert oracle_mod.cmt_exp(x) == oracle_mod.cmt_exp(x, ref=True)
        assert oracle_mod.cmt_sinhxoverx(x) == oracle_mod.cmt_sinhxoverx(x, ref=True)


def test_library_profiles_match_reference(oracle_mod):
    """The product's host table generator (mtmd_profile in libmtmd.so) gives
    the reference's float tables bit for bit."""
    from moltwister_amd.engine import profile
    for which, kind, prm, x0, x1 in PROFILES:
        a = profile(which, kind, prm, x0, x1)
        b = oracle_mod.profile(which, kind, prm, x0, x1, ref=True)
        for u, v in zip(a, b):
            assert np.array_equal(u, v, equal_nan=True)
