"""The oracle against committed golden fixtures (tests/golden/*.npz), written
from MolTwister's own CPU sources by tests/golden/make_golden.py. These run
everywhere, with or without the reference tree: bit-identical profiles, cell
lists, neighbour lists, per-atom forces/energies and NH-chain propagation."""
import glob
import os

import numpy as np
import pytest

from moltwister_amd import abi, config
from moltwister_amd.system import System

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(HERE, "*.npz")) if "profiles" not in p and "dcd_frames" not in p)


def load_case(name):
    z = np.load(os.path.join(HERE, f"{name}.npz"))
    nb = [(int(r[0]), int(r[1]), int(r[2]), tuple(r[3:])) for r in z["nonbonded"]]
    bff = [(int(r[0]), tuple(r[1:])) for r in z["bond_ff"]]
    aff = [(int(r[0]), tuple(r[1:])) for r in z["angle_ff"]]
    dff = [(int(r[0]), tuple(r[1:])) for r in z["dihedral_ff"]]
    s = System(pos=z["pos"], mom=z["mom"], mass=z["mass"], charge=z["charge"], types=z["types"],
               box=tuple(z["box"]), nonbonded=nb, bond_ff=bff, angle_ff=aff, dihedral_ff=dff,
               bonds=z["bonds"], angles=z["angles"], dihedrals=z["dihedrals"])
    cf = z["cfg"]
    c = config.defaults()
    c.cutoff_radius, c.neigh_shell = cf[0], cf[1]
    c.scale12, c.scale13, c.scale14, c.scale1N = cf[2], cf[3], cf[4], cf[5]
    c.timestep, c.temperature, c.temperature_relax = cf[6], cf[7], cf[8]
    c.temperature_nh_len, c.temperature_respa = int(cf[9]), int(cf[10])
    c.ensemble = abi.NVT
    return z, s, c


def test_fixtures_present():
    assert len(CASES) >= 5


@pytest.mark.parametrize("name", CASES)
def test_oracle_matches_golden(oracle_mod, name):
    z, s, c = load_case(name)
    o = oracle_mod.Oracle(s, c)
    cells = o.build_cells()
    assert np.array_equal(cells["nc"], z["nc"])
    assert np.array_equal(cells["ixyz"], z["ixyz"].astype(np.int32))
    assert np.array_equal(cells["first"], z["cell_first"])
    assert np.array_equal(cells["count"], z["cell_count"])
    assert np.array_equal(cells["list"], z["cell_list"])
    cnt, lst = o.build_neighbors()
    assert np.array_equal(cnt, z["ncount"])
    flat = np.concatenate([lst[k, :cnt[k]] for k in range(s.n_atoms)])
    assert np.array_equal(flat, z["nlist"])
    F, U = o.compute_forces()
    assert np.array_equal(F, z["F"])
    assert np.array_equal(U, z["U"])
    o.set_thermostat(z["nh_eta_in"], z["nh_peta_in"])
    o.nh_propagate()
    _, mom = o.get_state()
    assert np.array_equal(mom, z["nh_mom_out"])
    eta, pe = o.get_thermostat()
    assert np.array_equal(eta, z["nh_eta_out"]) and np.array_equal(pe, z["nh_peta_out"])


def test_profiles_match_golden(oracle_mod):
    z = np.load(os.path.join(HERE, "profiles.npz"))
    i = 0
    while f"p{i}_x" in z:
        a = z[f"p{i}_args"]
        which, kind = int(a[0]), int(a[1])
        prm = tuple(z[f"p{i}_prm"]) if f"p{i}_prm" in z else tuple(a[2:5])
        x0, x1 = float(np.float32(a[5])), float(np.float32(a[6]))
        x, f, e = oracle_mod.profile(which, kind, prm, x0, x1)
        assert np.array_equal(x, z[f"p{i}_x"]) and np.array_equal(f, z[f"p{i}_f"], equal_nan=True)
        assert np.array_equal(e, z[f"p{i}_e"], equal_nan=True)
        i += 1
    assert i >= 14


def test_library_profiles_match_golden():
    """libmtmd.so's own table generator (host code, loads without a GPU)."""
    from moltwister_amd.engine import profile
    z = np.load(os.path.join(HERE, "profiles.npz"))
    i = 0
    while f"p{i}_x" in z:
        a = z[f"p{i}_args"]
        prm = tuple(z[f"p{i}_prm"]) if f"p{i}_prm" in z else tuple(a[2:5])
        x, f, e = profile(int(a[0]), int(a[1]), prm, float(np.float32(a[5])), float(np.float32(a[6])))
        assert np.array_equal(x, z[f"p{i}_x"]) and np.array_equal(f, z[f"p{i}_f"], equal_nan=True)
        assert np.array_equal(e, z[f"p{i}_e"], equal_nan=True)
        i += 1


def as_profiles(s, rc):
    """The same system with every force field passed as its tabulated profile
    (MTMD_FF_PROFILE), the way a MolTwister binding hands over the tables its
    MDFF classes compute."""
    import copy
    from moltwister_amd.engine import profile
    t = copy.copy(s)
    rcf = float(np.float32(rc))
    t.nonbonded = [(a, b, abi.FF_PROFILE, np.stack(profile(0, k, p, 0.01, rcf))) for a, b, k, p in s.nonbonded]
    t.bond_ff = [(abi.FF_PROFILE, np.stack(profile(1, k, p, 0.01, rcf))) for k, p in s.bond_ff]
    two_pi = float(np.float32(2 * np.pi))
    t.angle_ff = [(abi.FF_PROFILE, np.stack(profile(2, k, p, 0.0, two_pi))) for k, p in s.angle_ff]
    t.dihedral_ff = [(abi.FF_PROFILE, np.stack(profile(3, k, p, 0.0, two_pi))) for k, p in s.dihedral_ff]
    return t


@pytest.mark.parametrize("name", ["water_300", "chains", "binary_lj"])
def test_oracle_profile_kind_matches_parameters(oracle_mod, name):
    """Force fields passed as profiles give the same forces, bit for bit."""
    from systems import make
    s, c = make(name)
    a = oracle_mod.Oracle(s, c).compute_forces()
    b = oracle_mod.Oracle(as_profiles(s, c.cutoff_radius), c).compute_forces()
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
