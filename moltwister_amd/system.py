"""Host-side description of an MD system, handed to libmtmd.so as mtmd_system.

Plays the role of the parts of CMolTwisterState the `moldyn` path reads
(MolTwisterState.h:43-101): atoms (name, position, mass, charge, mobility), the
MD force-field lists (`add mdnonbonded/mdbond/mdangle/mddihedral`) and the
periodic box. Atom-type indices follow the first-appearance order of the atom
names, as CMolTwisterState::searchForAtomTypes does (MolTwisterState.cpp:459-484).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import abi


def _ptr(a: np.ndarray | None, ctype):
    if a is None:
        return C.cast(None, C.POINTER(ctype))
    return a.ctypes.data_as(C.POINTER(ctype))


@dataclass
class System:
    pos: np.ndarray                      # (n, 3) float64, Angstrom
    mass: np.ndarray                     # (n,) g/mol
    charge: np.ndarray                   # (n,) e
    types: np.ndarray                    # (n,) int32 type indices
    box: tuple                           # (Lx, Ly, Lz)
    type_names: list = field(default_factory=list)
    nonbonded: list = field(default_factory=list)    # [(type_a, type_b, kind, params)]
    bond_ff: list = field(default_factory=list)      # [(kind, params)]
    angle_ff: list = field(default_factory=list)
    dihedral_ff: list = field(default_factory=list)
    bonds: np.ndarray | None = None      # (nb, 3) int32: a, b, ff
    angles: np.ndarray | None = None     # (na, 4) int32: a, centre, c, ff
    dihedrals: np.ndarray | None = None  # (nd, 5) int32
    mom: np.ndarray | None = None        # (n, 3) or None
    mobile: np.ndarray | None = None     # (n,) uint8 or None
    seed: int = 0

    def __post_init__(self) -> None:
        self.pos = np.ascontiguousarray(self.pos, dtype=np.float64).reshape(-1, 3)
        n = self.pos.shape[0]
        self.mass = np.ascontiguousarray(self.mass, dtype=np.float64).reshape(n)
        self.charge = np.ascontiguousarray(self.charge, dtype=np.float64).reshape(n)
        self.types = np.ascontiguousarray(self.types, dtype=np.int32).reshape(n)
        if self.mom is not None:
            self.mom = np.ascontiguousarray(self.mom, dtype=np.float64).reshape(n, 3)
        if self.mobile is not None:
            self.mobile = np.ascontiguousarray(self.mobile, dtype=np.uint8).reshape(n)
        for name, w in (("bonds", 3), ("angles", 4), ("dihedrals", 5)):
            a = getattr(self, name)
            a = np.zeros((0, w), np.int32) if a is None else np.ascontiguousarray(a, dtype=np.int32).reshape(-1, w)
            setattr(self, name, a)
        self.box = tuple(float(b) for b in self.box)

    @property
    def n_atoms(self) -> int:
        return self.pos.shape[0]

    @property
    def n_types(self) -> int:
        return int(self.types.max()) + 1 if self.n_atoms else 0

    def to_c(self) -> abi.MtmdSystem:
        """Build the mtmd_system view. The returned struct keeps references to
        every buffer it points into (attribute ``_keep``)."""
        s = abi.MtmdSystem()
        keep = []
        s.n_atoms = self.n_atoms
        s.n_types = max(self.n_types, max([max(a, b) + 1 for a, b, _, _ in self.nonbonded], default=0))
        s.pos = _ptr(self.pos, C.c_double)
        s.mom = _ptr(self.mom, C.c_double)
        s.mass = _ptr(self.mass, C.c_double)
        s.charge = _ptr(self.charge, C.c_double)
        s.type = _ptr(self.types, C.c_int32)
        s.mobile = _ptr(self.mobile, C.c_uint8)
        s.box[:] = self.box
        def profile_ptr(prm):
            # FF_PROFILE entries carry a (3, 1000) array x, -dU/dx, U instead of parameters
            arr = np.ascontiguousarray(np.asarray(prm, dtype=np.float32).reshape(3, -1))
            keep.append(arr)
            return arr.ctypes.data_as(C.POINTER(C.c_float))

        nb = (abi.MtmdNonbondedFF * max(1, len(self.nonbonded)))()
        for i, (a, b, kind, prm) in enumerate(self.nonbonded):
            nb[i].type_a, nb[i].type_b, nb[i].kind = a, b, kind
            if kind == abi.FF_PROFILE:
                nb[i].profile = profile_ptr(prm)
                continue
            p = list(prm) + [0.0] * (4 - len(prm))
            if kind in (abi.NB_LJ, abi.NB_LJ1208) and len(prm) < 3:
                p[2] = 1.0   # alpha defaults to 1 (MolTwisterMDFFNonBonded_LJ.h:29)
            nb[i].p[:] = p
        s.n_nonbonded, s.nonbonded = len(self.nonbonded), C.cast(nb, C.POINTER(abi.MtmdNonbondedFF))
        keep.append(nb)
        for lst, nfield, pfield in ((self.bond_ff, "n_bond_ff", "bond_ff"),
                                    (self.angle_ff, "n_angle_ff", "angle_ff"),
                                    (self.dihedral_ff, "n_dihedral_ff", "dihedral_ff")):
            arr = (abi.MtmdBondedFF * max(1, len(lst)))()
            for i, (kind, prm) in enumerate(lst):
                arr[i].kind = kind
                if kind == abi.FF_PROFILE:
                    arr[i].profile = profile_ptr(prm)
                else:
                    arr[i].p[:] = list(prm) + [0.0] * (abi.BONDED_PARAMS - len(prm))
            setattr(s, nfield, len(lst))
            setattr(s, pfield, C.cast(arr, C.POINTER(abi.MtmdBondedFF)))
            keep.append(arr)
        s.n_bonds, s.bonds = len(self.bonds), _ptr(self.bonds, C.c_int32)
        s.n_angles, s.angles = len(self.angles), _ptr(self.angles, C.c_int32)
        s.n_dihedrals, s.dihedrals = len(self.dihedrals), _ptr(self.dihedrals, C.c_int32)
        s.seed = int(self.seed) & 0xFFFFFFFFFFFFFFFF
        if self.type_names:
            names = (C.c_char_p * len(self.type_names))(*[str(x).encode() for x in self.type_names])
            s.type_names = C.cast(names, C.POINTER(C.c_char_p))
            keep.append(names)
        keep.extend([self.pos, self.mom, self.mass, self.charge, self.types, self.mobile,
                     self.bonds, self.angles, self.dihedrals])
        s._keep = keep
        return s


def type_indices(names) -> tuple[np.ndarray, list]:
    """Type index per atom in first-appearance order of the names
    (CMolTwisterState::searchForAtomTypes, MolTwisterState.cpp:459-484)."""
    order: list = []
    idx = np.empty(len(names), np.int32)
    lookup: dict = {}
    for i, nm in enumerate(names):
        if nm not in lookup:
            lookup[nm] = len(order)
            order.append(nm)
        idx[i] = lookup[nm]
    return idx, order
