"""The C ABI (include/mtmd.h) on the host: libmtmd.so loads, exports every
declared entry point, and its configuration surface behaves like the
reference's `moldyn cfg` (Cmd/MolTwisterCmdMolDyn/Config/MolDynConfig.cpp).
No GPU needed."""
import ctypes as C
import os
import re

import pytest

from moltwister_amd import abi, config

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "mtmd.h")).read()
    return sorted(set(re.findall(r"\b(mtmd_[a-z_]+)\s*\(", text)))


def test_library_loads_and_exports_every_declared_symbol():
    lib = abi.load_library()
    syms = header_symbols()
    assert len(syms) >= 20
    for name in syms:
        assert hasattr(lib, name), name
    assert sorted(abi.EXPORTED) == syms
    assert lib.mtmd_abi_version() == 4


def test_defaults_match_reference():
    lib = abi.load_library()
    c = abi.MtmdConfig()
    lib.mtmd_config_defaults(C.byref(c))
    # CMolDynConfig::resetToDefaults (MolDynConfig.cpp:289-328)
    assert c.timestep == 1.0 and c.out_stride == 100 and c.ensemble == abi.NVT
    assert c.temperature == 298.0 and c.temperature_relax == 20.0
    assert c.temperature_nh_len == 4 and c.temperature_respa == 4
    assert c.pressure == 1.0 and c.pressure_relax == 5000.0
    assert c.cutoff_radius == 10.0 and c.neigh_shell == 2.0 and c.cutoff_force == 1000.0
    assert c.n_steps == 50000
    assert (c.scale12, c.scale13, c.scale14, c.scale1N) == (0.0, 0.0, 0.5, 0.0)
    assert c.opt_learning_rate == 1e-6 and c.opt_max_steps == 5000 and c.opt_accuracy == 0.0
    assert c.info_file == b"out.txt"
    py = config.defaults()
    for name, _ in abi.MtmdConfig._fields_:
        if name == "reserved":
            continue
        assert getattr(py, name) == getattr(c, name), name


@pytest.mark.parametrize("key,value,field,expect", [
    ("timestep", "0.5", "timestep", 0.5),
    ("timesteps", "1000", "n_steps", 1000),
    ("outstride", "10", "out_stride", 10),
    ("temperature", "86.596", "temperature", 86.596),
    ("cutoffradius", "2.5", "cutoff_radius", 2.5),
    ("neighshell", "0.3", "neigh_shell", 0.3),
    ("scale14", "0.75", "scale14", 0.75),
    ("temperaturenhlen", "6", "temperature_nh_len", 6),
    ("ensemble", "nve", "ensemble", abi.NVE),
    ("ensemble", "npt", "ensemble", abi.NPT),
    ("engine", "mixed", "engine", abi.ENGINE_MIXED),
])
def test_config_set(key, value, field, expect):
    lib = abi.load_library()
    c = abi.MtmdConfig()
    lib.mtmd_config_defaults(C.byref(c))
    err = C.create_string_buffer(256)
    assert lib.mtmd_config_set(C.byref(c), key.encode(), value.encode(), err, 256) == 0
    assert getattr(c, field) == pytest.approx(expect)
    m = config.MolDynConfig()
    assert m.set(key, value) == ""
    assert getattr(m.cfg, field) == pytest.approx(expect)


@pytest.mark.parametrize("key,value,msg", [
    ("ensemble", "nvx", "Error: unknown ensemble!"),
    ("bogus", "1", "Error: unknown parameter!"),
    ("includedcdfile", "maybe", "Error: 'includedcdfile' parameter should be either 'yes' or 'no'!"),
])
def test_config_set_errors(key, value, msg):
    lib = abi.load_library()
    c = abi.MtmdConfig()
    lib.mtmd_config_defaults(C.byref(c))
    err = C.create_string_buffer(256)
    assert lib.mtmd_config_set(C.byref(c), key.encode(), value.encode(), err, 256) == abi.MTMD_EINVAL
    assert err.value.decode() == msg
    assert config.MolDynConfig().set(key, value) == msg


def test_atoi_atof_semantics():
    m = config.MolDynConfig()
    assert m.set("timesteps", "12abc") == "" and m.cfg.n_steps == 12
    assert m.set("timestep", "x") == "" and m.cfg.timestep == 0.0


def test_status_strings():
    lib = abi.load_library()
    for code in (0, -1, -2, -3, -4, -5, -6, -7):
        assert lib.mtmd_status_string(code)


def test_create_rejects_bad_input_before_touching_the_gpu():
    """Invalid arguments are refused with MTMD_EINVAL/ENODEV and a message."""
    lib = abi.load_library()
    h = C.c_void_p()
    assert lib.mtmd_create(None, None, C.byref(h)) == abi.MTMD_EINVAL
    assert lib.mtmd_last_error(None)
    assert lib.mtmd_step(None, 1) == abi.MTMD_EINVAL
    assert lib.mtmd_compute_forces(None, None, None, None) == abi.MTMD_EINVAL


def test_profile_rejects_unknown_kind():
    lib = abi.load_library()
    import numpy as np
    p = np.zeros(4)
    assert lib.mtmd_profile(0, 99, p.ctypes.data_as(C.POINTER(C.c_double)), 0.01, 2.5, 10, None, None, None) \
        == abi.MTMD_EINVAL
