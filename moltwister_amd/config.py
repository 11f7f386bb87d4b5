"""`moldyn cfg` semantics: SMolDynConfigStruct keys, defaults and setters.

Mirrors CMolDynConfig (Cmd/MolTwisterCmdMolDyn/Config/MolDynConfig.cpp):
keys (:28-41), ``set`` (:95-287) returning "" or the reference's "Error: ..."
strings, defaults (:289-328). The engine options (``engine``, ``coulomb``,
``skin``, ``blocksize``) have no counterpart in the reference.
"""
from __future__ import annotations

from . import abi

KEYWORDS = [
    "timestep", "timesteps", "outstride", "ensemble", "NVT", "NPT",
    "temperature", "temperaturerelax", "temperaturenhlen", "temperaturerespa",
    "pressure", "pressurerelax", "pressurenhlen", "pressurerespa",
    "cutoffradius", "neighshell", "cutoffforce",
    "infofile", "xyzfile", "dcdfile", "xtcfile", "pdistrfile", "vdistrfile",
    "includexyzfile", "includedcdfile", "includextcfile", "xtcprecision",
    "includepdistrfile", "includevdistrfile", "maxpdistr", "maxvdistr", "verboseoutput",
    "scale12", "scale13", "scale14", "scale1N", "optmaxsteps", "optlearningrate",
    "optaccuracy",
]

# key -> (field, parser)
_NUMERIC = {
    "timesteps": ("n_steps", int),
    "timestep": ("timestep", float),
    "outstride": ("out_stride", int),
    "temperature": ("temperature", float),
    "temperaturerelax": ("temperature_relax", float),
    "temperaturenhlen": ("temperature_nh_len", int),
    "temperaturerespa": ("temperature_respa", int),
    "pressure": ("pressure", float),
    "pressurerelax": ("pressure_relax", float),
    "pressurenhlen": ("pressure_nh_len", int),
    "pressurerespa": ("pressure_respa", int),
    "cutoffradius": ("cutoff_radius", float),
    "neighshell": ("neigh_shell", float),
    "cutoffforce": ("cutoff_force", float),
    "scale12": ("scale12", float),
    "scale13": ("scale13", float),
    "scale14": ("scale14", float),
    "scale1N": ("scale1N", float),
    "optlearningrate": ("opt_learning_rate", float),
    "optmaxsteps": ("opt_max_steps", int),
    "optaccuracy": ("opt_accuracy", float),
    # engine options
    "skin": ("skin", float),
    "blocksize": ("block_size", int),
}
# key -> mtmd_config field
_YESNO = {"includexyzfile": "include_xyz", "includedcdfile": "include_dcd", "includextcfile": "include_xtc",
          "includepdistrfile": "include_pdistr", "includevdistrfile": "include_vdistr", "verboseoutput": "verbose"}
_FILES = {"infofile": "info_file", "xyzfile": "xyz_file", "dcdfile": "dcd_file", "xtcfile": "xtc_file",
          "pdistrfile": "pdistr_file", "vdistrfile": "vdistr_file"}


def _atoi(s: str) -> int:
    """C atoi: leading integer prefix, 0 when none."""
    s = s.strip()
    i, sign = 0, 1
    if s[:1] in "+-":
        sign = -1 if s[0] == "-" else 1
        i = 1
    j = i
    while j < len(s) and s[j].isdigit():
        j += 1
    return sign * int(s[i:j]) if j > i else 0


def _atof(s: str) -> float:
    """C atof: longest float prefix, 0.0 when none."""
    s = s.strip()
    for end in range(len(s), 0, -1):
        try:
            return float(s[:end])
        except ValueError:
            continue
    return 0.0


def defaults() -> abi.MtmdConfig:
    """CMolDynConfig::resetToDefaults (MolDynConfig.cpp:289-328)."""
    c = abi.MtmdConfig()
    c.timestep = 1.0
    c.out_stride = 100
    c.ensemble = abi.NVT
    c.temperature = 298.0
    c.temperature_relax = 20.0 * c.timestep
    c.temperature_nh_len = 4
    c.temperature_respa = 4
    c.pressure = 1.0
    c.pressure_relax = 5000.0 * c.timestep
    c.pressure_nh_len = 4
    c.pressure_respa = 4
    c.cutoff_radius = 10.0
    c.neigh_shell = 2.0
    c.cutoff_force = 1000.0
    c.n_steps = 50000
    c.info_file = b"out.txt"
    c.xyz_file = b"traj.xyz"
    c.dcd_file = b"traj.dcd"
    c.xtc_file = b"traj.xtc"
    c.pdistr_file = b"distr"
    c.vdistr_file = b"distr"
    c.include_xyz = 0
    c.include_dcd = 1
    c.include_xtc = 0
    c.xtc_precision = 1000.0
    c.include_pdistr = 0
    c.include_vdistr = 0
    c.max_pdistr = 0.3
    c.max_vdistr = 100.0e3
    c.verbose = 0
    c.scale12 = 0.0
    c.scale13 = 0.0
    c.scale14 = 0.5
    c.scale1N = 0.0
    c.opt_learning_rate = 1e-6
    c.opt_max_steps = 5000
    c.opt_accuracy = 0.0
    c.engine = abi.ENGINE_EXACT
    c.coulomb = abi.COUL_AUTO
    c.skin = 0.3
    c.block_size = 0
    return c


class MolDynConfig:
    """`moldyn cfg get/set/resettodefaults` over an mtmd_config."""

    def __init__(self) -> None:
        self.cfg = defaults()

    def reset_to_defaults(self) -> None:
        self.__init__()

    def set(self, parameter: str, value: str) -> str:
        """CMolDynConfig::set (MolDynConfig.cpp:95-287): "" on success."""
        c = self.cfg
        if parameter in _NUMERIC:
            field, kind = _NUMERIC[parameter]
            setattr(c, field, _atoi(value) if kind is int else _atof(value))
        elif parameter == "ensemble":
            m = {"nvt": abi.NVT, "npt": abi.NPT, "nve": abi.NVE}
            if value not in m:
                return "Error: unknown ensemble!"
            c.ensemble = m[value]
        elif parameter in _FILES:
            setattr(c, _FILES[parameter], value.encode()[:255])
        elif parameter in _YESNO:
            if value == "yes":
                setattr(c, _YESNO[parameter], 1)
            elif value == "no":
                setattr(c, _YESNO[parameter], 0)
            else:
                return f"Error: '{parameter}' parameter should be either 'yes' or 'no'!"
        elif parameter == "xtcprecision":
            c.xtc_precision = _atof(value)
        elif parameter == "maxpdistr":
            c.max_pdistr = _atof(value)
        elif parameter == "maxvdistr":
            c.max_vdistr = _atof(value)
        elif parameter == "engine":
            m = {"exact": abi.ENGINE_EXACT, "fp64": abi.ENGINE_FP64, "mixed": abi.ENGINE_MIXED}
            if value not in m:
                return "Error: unknown engine!"
            c.engine = m[value]
        elif parameter == "coulomb":
            m = {"auto": abi.COUL_AUTO, "allpairs": abi.COUL_ALLPAIRS, "cutoff": abi.COUL_CUTOFF}
            if value not in m:
                return "Error: unknown coulomb mode!"
            c.coulomb = m[value]
        else:
            return "Error: unknown parameter!"
        return ""
