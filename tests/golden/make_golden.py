"""Write the golden fixtures in this directory from MolTwister's own CPU
sources (oracle/_ref/libmtref.so, built from the MolTwister sources by
oracle/Makefile). Inputs and outputs are stored together, so the fixtures do
not depend on the generators that made the inputs.

  python tests/golden/make_golden.py [--profiles-only]
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))

from moltwister_amd import abi  # noqa: E402
from oracle import pyoracle  # noqa: E402
from systems import make  # noqa: E402

CASES = ["fcc_2cells", "binary_lj", "water_300", "chains", "buckingham"]


def system_arrays(s, c):
    d = dict(pos=s.pos, mom=s.mom, mass=s.mass, charge=s.charge, types=s.types, box=np.array(s.box),
             bonds=s.bonds, angles=s.angles, dihedrals=s.dihedrals,
             nonbonded=np.array([[a, b, k, *(list(p) + [0.0] * (4 - len(p)))] for a, b, k, p in s.nonbonded], float),
             bond_ff=np.array([[k, *(list(p) + [0.0] * (4 - len(p)))] for k, p in s.bond_ff] or np.zeros((0, 5)), float),
             angle_ff=np.array([[k, *(list(p) + [0.0] * (4 - len(p)))] for k, p in s.angle_ff] or np.zeros((0, 5)), float),
             dihedral_ff=np.array([[k, *(list(p) + [0.0] * (4 - len(p)))] for k, p in s.dihedral_ff] or np.zeros((0, 5)), float),
             cfg=np.array([c.cutoff_radius, c.neigh_shell, c.scale12, c.scale13, c.scale14, c.scale1N, c.timestep,
                           c.temperature, c.temperature_relax, c.temperature_nh_len, c.temperature_respa]))
    return d


def main():
    pyoracle.build(with_ref=True)
    if "--profiles-only" in sys.argv:
        return profiles_and_dcd()
    for name in CASES:
        s, c = make(name, ensemble=abi.NVT)
        r = pyoracle.Reference(s, c)
        cells = r.build_cells()
        cnt, lst = r.build_neighbors()
        F, U = r.compute_forces()
        flat = np.concatenate([lst[k, :cnt[k]] for k in range(s.n_atoms)]).astype(np.int32)
        eta = np.array([0.1, 0.2, 0.3, 0.4])
        pe = np.array([16.0, -3.0, 2.0, 0.5])
        r.set_thermostat(eta, pe)
        mom_nh = r.nh_propagate()
        eta2, pe2 = r.get_thermostat()
        out = system_arrays(s, c)
        out.update(nc=cells["nc"], ixyz=cells["ixyz"].astype(np.int16), cell_first=cells["first"],
                   cell_count=cells["count"], cell_list=cells["list"], ncount=cnt, nlist=flat, F=F, U=U,
                   nh_eta_in=eta, nh_peta_in=pe, nh_mom_out=mom_nh, nh_eta_out=eta2, nh_peta_out=pe2)
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **out)
        print(name, s.n_atoms, os.path.getsize(os.path.join(HERE, f"{name}.npz")))
    profiles_and_dcd()


def profiles_and_dcd():
    prof = {}
    from test_oracle_ref import PROFILES
    for i, (which, kind, prm, x0, x1) in enumerate(PROFILES):
        x, f, e = pyoracle.profile(which, kind, prm, x0, x1, ref=True)
        prof[f"p{i}_x"], prof[f"p{i}_f"], prof[f"p{i}_e"] = x, f, e
        prof[f"p{i}_args"] = np.array([which, kind, *(list(prm[:3]) + [0.0] * (3 - len(prm[:3]))), x0, x1], float)
        prof[f"p{i}_prm"] = np.array(prm, float)
    np.savez_compressed(os.path.join(HERE, "profiles.npz"), **prof)
    # two DCD frames written by MolTwister's CDCDFile (header + records)
    import tempfile
    rng = np.random.default_rng(5)
    frames = rng.uniform(-6.0, 6.0, (2, 11, 3))
    box = np.array([12.75, 13.2, 9.99])
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "ref.dcd")
        for f in frames:
            pyoracle.ref_dcd_append(path, 1000, 100, 0.5, f, box)
        data = np.frombuffer(open(path, "rb").read(), dtype=np.uint8)
    np.savez_compressed(os.path.join(HERE, "dcd_frames.npz"), frames=frames, box=box,
                        args=np.array([1000, 100, 0.5]), dcd=data)


if __name__ == "__main__":
    main()
