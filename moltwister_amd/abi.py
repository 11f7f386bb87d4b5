"""ctypes view of the C ABI declared in include/mtmd.h.

The structures mirror ``mtmd_config``, ``mtmd_system``, ``mtmd_thermo`` and the
force-field records field for field. ``load_library()`` opens the in-tree
``libmtmd.so`` and raises if it is missing: there is no Python fallback for the
engine.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

MTMD_NUM_POINTS = 1000
MTMD_MAX_FF_PER_PAIR = 5

MTMD_OK = 0
MTMD_EINVAL = -1
MTMD_EHIP = -2
MTMD_ENEIGH = -3
MTMD_ENAN = -4
MTMD_EBOX = -5
MTMD_ENOMEM = -6
MTMD_ENODEV = -7

NVT, NPT, NVE = 0, 1, 2
ENGINE_EXACT, ENGINE_FP64, ENGINE_MIXED = 0, 1, 2
COUL_AUTO, COUL_ALLPAIRS, COUL_CUTOFF = 0, 1, 2

NB_LJ, NB_BUCK, NB_LJ1208, NB_LJBUCK = 1, 2, 3, 4
FF_PROFILE = 100          # force field given as its (x, -dU/dx, U) profile
BOND_HARM, BOND_MORSE = 1, 2
ANGLE_HARM, ANGLE_CLASS2 = 1, 2
DIH_HARM, DIH_FOURIER4T = 1, 2
BONDED_PARAMS = 12


class MtmdConfig(C.Structure):
    _fields_ = [
        ("ensemble", C.c_int32),
        ("temperature", C.c_double),
        ("temperature_relax", C.c_double),
        ("temperature_nh_len", C.c_int32),
        ("temperature_respa", C.c_int32),
        ("pressure", C.c_double),
        ("pressure_relax", C.c_double),
        ("pressure_nh_len", C.c_int32),
        ("pressure_respa", C.c_int32),
        ("cutoff_radius", C.c_double),
        ("neigh_shell", C.c_double),
        ("cutoff_force", C.c_double),
        ("n_steps", C.c_int32),
        ("timestep", C.c_double),
        ("out_stride", C.c_int32),
        ("scale12", C.c_double),
        ("scale13", C.c_double),
        ("scale14", C.c_double),
        ("scale1N", C.c_double),
        ("opt_learning_rate", C.c_double),
        ("opt_max_steps", C.c_int32),
        ("opt_accuracy", C.c_double),
        ("verbose", C.c_int32),
        ("info_file", C.c_char * 256),
        ("xyz_file", C.c_char * 256),
        ("dcd_file", C.c_char * 256),
        ("xtc_file", C.c_char * 256),
        ("pdistr_file", C.c_char * 256),
        ("vdistr_file", C.c_char * 256),
        ("include_xyz", C.c_int32),
        ("include_dcd", C.c_int32),
        ("include_xtc", C.c_int32),
        ("xtc_precision", C.c_float),
        ("include_pdistr", C.c_int32),
        ("include_vdistr", C.c_int32),
        ("max_pdistr", C.c_double),
        ("max_vdistr", C.c_double),
        ("engine", C.c_int32),
        ("coulomb", C.c_int32),
        ("skin", C.c_double),
        ("block_size", C.c_int32),
        ("reserved", C.c_int32 * 7),
    ]


class MtmdNonbondedFF(C.Structure):
    _fields_ = [("type_a", C.c_int32), ("type_b", C.c_int32), ("kind", C.c_int32),
                ("pad", C.c_int32), ("p", C.c_double * 4), ("profile", C.POINTER(C.c_float))]


class MtmdBondedFF(C.Structure):
    _fields_ = [("kind", C.c_int32), ("pad", C.c_int32), ("p", C.c_double * BONDED_PARAMS),
                ("profile", C.POINTER(C.c_float))]


_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class MtmdSystem(C.Structure):
    _fields_ = [
        ("n_atoms", C.c_int32),
        ("n_types", C.c_int32),
        ("pos", _dp),
        ("mom", _dp),
        ("mass", _dp),
        ("charge", _dp),
        ("type", _ip),
        ("mobile", C.POINTER(C.c_uint8)),
        ("box", C.c_double * 3),
        ("n_nonbonded", C.c_int32),
        ("nonbonded", C.POINTER(MtmdNonbondedFF)),
        ("n_bond_ff", C.c_int32),
        ("bond_ff", C.POINTER(MtmdBondedFF)),
        ("n_angle_ff", C.c_int32),
        ("angle_ff", C.POINTER(MtmdBondedFF)),
        ("n_dihedral_ff", C.c_int32),
        ("dihedral_ff", C.POINTER(MtmdBondedFF)),
        ("n_bonds", C.c_int32),
        ("bonds", _ip),
        ("n_angles", C.c_int32),
        ("angles", _ip),
        ("n_dihedrals", C.c_int32),
        ("dihedrals", _ip),
        ("seed", C.c_uint64),
        ("type_names", C.POINTER(C.c_char_p)),
    ]


class MtmdTopoRule(C.Structure):
    _fields_ = [("types", C.c_int32 * 4), ("R0", C.c_double), ("crit", C.c_int32), ("reserved", C.c_int32)]


# detection criteria of a topology rule (CMDFFBond::EDetCrit)
CRIT_ALL_R, CRIT_MOL_R, CRIT_VISIBLE, CRIT_14 = 0, 1, 2, 3


def topo_rules(rules):
    """An mtmd_topo_rule array from [(types tuple, R0[, crit])]; crit defaults
    to CRIT_ALL_R."""
    arr = (MtmdTopoRule * max(1, len(rules)))()
    for i, rule in enumerate(rules):
        t, r0 = rule[0], rule[1]
        arr[i].types[:] = list(t) + [0] * (4 - len(t))
        arr[i].R0 = r0
        arr[i].crit = rule[2] if len(rule) > 2 else CRIT_ALL_R
    return arr


def bond_graph(graph, n):
    """(bond_start [n + 1], bond_dest) int32 arrays of a bond graph given as one
    list of bond destinations per atom, in the atom's own bond order; (None,
    None) for no graph."""
    if graph is None:
        return None, None
    if len(graph) != n:
        raise ValueError("the bond graph needs one list of destinations per atom")
    start = np.zeros(n + 1, np.int32)
    start[1:] = np.cumsum([len(g) for g in graph])
    dest = np.array([d for g in graph for d in g], np.int32)
    if dest.size and (dest.min() < 0 or dest.max() >= n):
        raise ValueError("bond destination out of range")
    return start, np.ascontiguousarray(dest if dest.size else np.zeros(1, np.int32))


class MtmdThermo(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("T", "P", "V", "U", "K", "Etot", "sum_p2m", "sum_fr", "U_cut")]


PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libmtmd.so")

# every symbol include/mtmd.h declares
EXPORTED = [
    "mtmd_abi_version", "mtmd_last_error", "mtmd_status_string", "mtmd_config_defaults",
    "mtmd_config_set", "mtmd_profile", "mtmd_create", "mtmd_destroy", "mtmd_set_state",
    "mtmd_get_state", "mtmd_set_thermostat", "mtmd_get_thermostat", "mtmd_build_cells",
    "mtmd_build_neighbors", "mtmd_compute_forces", "mtmd_step", "mtmd_get_thermo", "mtmd_get_thermo_noalloc",
    "mtmd_get_forces", "mtmd_optimize", "mtmd_get_stats", "mtmd_get_skin_lists", "mtmd_synchronize", "mtmd_run",
    "mtmd_run_optimize", "mtmd_bench_kernel", "mtmd_append_xyz", "mtmd_append_dcd",
    "mtmd_set_barostat", "mtmd_get_barostat", "mtmd_find_topology", "mtmd_topology_last_error",
    "mtmd_batch_create", "mtmd_batch_destroy", "mtmd_batch_last_error", "mtmd_batch_step",
    "mtmd_batch_step_separate", "mtmd_batch_synchronize", "mtmd_batch_get_thermo", "mtmd_batch_get_state",
    "mtmd_batch_get_forces", "mtmd_batch_autotune", "mtmd_batch_info", "mtmd_batch_tuning",
    "mtmd_reference_draws",
    "mtmd_fm_create", "mtmd_fm_destroy", "mtmd_fm_last_error", "mtmd_fm_num_params", "mtmd_fm_set_target",
    "mtmd_fm_target_from_params", "mtmd_fm_forces", "mtmd_fm_loss_grad",
]

_lib = None


def load_library(path: str | None = None) -> C.CDLL:
    """Open libmtmd.so (built in-tree by ``moltwister_amd.build``). Raises
    FileNotFoundError when it has not been built: the engine has no fallback."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise FileNotFoundError(
            f"{p} is missing: build the HIP engine first (python -m moltwister_amd.build)")
    lib = C.CDLL(p)
    vp = C.c_void_p
    sig = {
        "mtmd_abi_version": (C.c_int, []),
        "mtmd_last_error": (C.c_char_p, [vp]),
        "mtmd_status_string": (C.c_char_p, [C.c_int]),
        "mtmd_config_defaults": (None, [C.POINTER(MtmdConfig)]),
        "mtmd_config_set": (C.c_int, [C.POINTER(MtmdConfig), C.c_char_p, C.c_char_p, C.c_char_p, C.c_size_t]),
        "mtmd_profile": (C.c_int, [C.c_int, C.c_int, _dp, C.c_float, C.c_float, C.c_int,
                                   C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float)]),
        "mtmd_create": (C.c_int, [C.POINTER(MtmdSystem), C.POINTER(MtmdConfig), C.POINTER(vp)]),
        "mtmd_destroy": (None, [vp]),
        "mtmd_set_state": (C.c_int, [vp, _dp, _dp]),
        "mtmd_get_state": (C.c_int, [vp, _dp, _dp]),
        "mtmd_set_thermostat": (C.c_int, [vp, _dp, _dp]),
        "mtmd_get_thermostat": (C.c_int, [vp, _dp, _dp]),
        "mtmd_build_cells": (C.c_int, [vp, _ip, _ip, _ip, _ip, _ip]),
        "mtmd_build_neighbors": (C.c_int, [vp, C.c_int32, _ip, _ip, _ip]),
        "mtmd_compute_forces": (C.c_int, [vp, _dp, _dp, C.POINTER(C.c_float)]),
        "mtmd_step": (C.c_int, [vp, C.c_int32]),
        "mtmd_get_thermo": (C.c_int, [vp, C.POINTER(MtmdThermo)]),
        "mtmd_get_thermo_noalloc": (C.c_int, [vp, C.POINTER(MtmdThermo)]),
        "mtmd_get_forces": (C.c_int, [vp, _dp, C.POINTER(C.c_float)]),
        "mtmd_optimize": (C.c_int, [vp, C.c_int32, C.c_double, C.c_double, _ip]),
        "mtmd_get_stats": (C.c_int, [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64), _ip]),
        "mtmd_get_skin_lists": (C.c_int, [vp, C.c_int32, _ip, _ip, _ip]),
        "mtmd_synchronize": (C.c_int, [vp]),
        "mtmd_run": (C.c_int, [C.POINTER(MtmdConfig), C.POINTER(MtmdSystem), vp]),
        "mtmd_run_optimize": (C.c_int, [C.POINTER(MtmdConfig), C.POINTER(MtmdSystem), vp]),
        "mtmd_bench_kernel": (C.c_int, [vp, C.c_int, C.c_int, C.POINTER(C.c_double)]),
        "mtmd_append_xyz": (C.c_int, [vp, C.c_char_p, C.c_int32]),
        "mtmd_set_barostat": (C.c_int, [vp, C.c_double, C.c_double, _dp, _dp]),
        "mtmd_find_topology": (C.c_int, [C.c_int32, C.c_int32, _dp, _ip, _dp, _ip, _ip, C.c_int32, C.POINTER(MtmdTopoRule),
                                         C.c_int32, _ip, _ip]),
        "mtmd_topology_last_error": (C.c_char_p, []),
        "mtmd_get_barostat": (C.c_int, [vp, _dp, _dp, _dp, _dp, _dp]),
        "mtmd_append_dcd": (C.c_int, [vp, C.c_char_p, C.c_int32, C.c_int32]),
        "mtmd_batch_create": (C.c_int, [C.c_int, C.POINTER(MtmdSystem), C.POINTER(MtmdConfig),
                                        C.POINTER(C.c_void_p)]),
        "mtmd_batch_destroy": (None, [vp]),
        "mtmd_batch_last_error": (C.c_char_p, [vp]),
        "mtmd_batch_step": (C.c_int, [vp, C.c_int32]),
        "mtmd_batch_step_separate": (C.c_int, [vp, C.c_int32]),
        "mtmd_batch_synchronize": (C.c_int, [vp]),
        "mtmd_batch_get_thermo": (C.c_int, [vp, C.POINTER(MtmdThermo)]),
        "mtmd_batch_get_state": (C.c_int, [vp, C.c_int32, _dp, _dp]),
        "mtmd_batch_get_forces": (C.c_int, [vp, C.c_int32, _dp, C.POINTER(C.c_float)]),
        "mtmd_batch_autotune": (C.c_int, [vp, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
        "mtmd_batch_info": (C.c_int, [vp, _ip, _ip, _ip, _ip]),
        "mtmd_batch_tuning": (C.c_int, [vp, _ip, _ip, _ip, _ip]),
        "mtmd_reference_draws": (C.c_int, [C.POINTER(MtmdSystem), C.POINTER(MtmdConfig), C.c_uint32, _dp, _dp, _dp,
                                           _dp, _dp, _dp]),
        "mtmd_fm_create": (C.c_int, [C.POINTER(MtmdSystem), C.c_double, C.c_int, _dp, _dp, C.POINTER(C.c_void_p)]),
        "mtmd_fm_destroy": (None, [vp]),
        "mtmd_fm_last_error": (C.c_char_p, [vp]),
        "mtmd_fm_num_params": (C.c_int, [vp]),
        "mtmd_fm_set_target": (C.c_int, [vp, _dp]),
        "mtmd_fm_target_from_params": (C.c_int, [vp, _dp]),
        "mtmd_fm_forces": (C.c_int, [vp, _dp, C.c_int32, C.c_int32, _dp]),
        "mtmd_fm_loss_grad": (C.c_int, [vp, _dp, C.c_int32, _dp, _dp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib
