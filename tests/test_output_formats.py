"""Output files of `moldyn run` on the host (no GPU): the DCD fixture written
by MolTwister's own CDCDFile (Utilities/DCDFile.cpp:137-246,670-729), read
back with an independent CHARMM-DCD parser, and the configuration keys that
name the output files (MolDynConfig.cpp:163-243,307-319)."""
import ctypes as C
import os
import struct

import numpy as np
import pytest

from moltwister_amd import abi, config

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "dcd_frames.npz")


def parse_dcd(data: bytes):
    """CHARMM DCD reader (Fortran unformatted records, little-endian)."""
    off = 0

    def record():
        nonlocal off
        (n,) = struct.unpack_from("<I", data, off)
        body = data[off + 4:off + 4 + n]
        (n2,) = struct.unpack_from("<I", data, off + 4 + n)
        assert n == n2, "record markers differ"
        off += 8 + n
        return body

    hdr = record()
    assert len(hdr) == 84 and hdr[:4] == b"CORD"
    icntrl = struct.unpack_from("<20i", hdr, 4)
    delta = struct.unpack_from("<f", hdr, 4 + 9 * 4)[0]
    titles = record()
    (ntitle,) = struct.unpack_from("<i", titles, 0)
    title = [titles[4 + 80 * i:4 + 80 * (i + 1)] for i in range(ntitle)]
    (natoms,) = struct.unpack("<i", record())
    frames, cells = [], []
    while off < len(data):
        cells.append(np.frombuffer(record(), "<f8"))
        xyz = [np.frombuffer(record(), "<f4") for _ in range(3)]
        frames.append(np.stack(xyz, 1))
    return dict(nset=icntrl[0], istart=icntrl[1], nsavc=icntrl[2], delta=delta, has_cell=icntrl[10],
                version=icntrl[19], titles=title, natoms=natoms, frames=frames, cells=cells)


def test_golden_dcd_layout():
    g = np.load(GOLD)
    d = parse_dcd(bytes(g["dcd"]))
    n_steps, stride, dt = g["args"]
    assert d["nset"] == int(n_steps) // int(stride) and d["istart"] == 0 and d["nsavc"] == int(stride)
    assert d["delta"] == np.float32(dt) and d["has_cell"] == 1 and d["version"] == 1
    assert d["titles"][0].rstrip(b"\0") == b"Written by MolTwister" and d["titles"][1].rstrip(b"\0") == b"---"
    assert d["natoms"] == g["frames"].shape[1] and len(d["frames"]) == g["frames"].shape[0]
    for f, ref in zip(d["frames"], g["frames"]):
        assert np.array_equal(f, ref.astype(np.float32))
    # unit cell: widths narrowed to int, angles 0, the reference's slot order
    # (boxX, angX, boxY, angY, angZ, boxZ)
    bx = np.trunc(g["box"])
    for c in d["cells"]:
        assert np.array_equal(c, [bx[0], 0.0, bx[1], 0.0, 0.0, bx[2]])


def test_golden_dcd_matches_reference_build(tmp_path):
    from oracle import pyoracle
    if not pyoracle.ref_available():
        pytest.skip("MolTwister reference build (oracle/_ref) not present")
    g = np.load(GOLD)
    path = str(tmp_path / "r.dcd")
    for f in g["frames"]:
        pyoracle.ref_dcd_append(path, int(g["args"][0]), int(g["args"][1]), float(g["args"][2]), f, g["box"])
    assert open(path, "rb").read() == bytes(g["dcd"])


def test_output_keys_defaults_and_set():
    lib = abi.load_library()
    c = abi.MtmdConfig()
    lib.mtmd_config_defaults(C.byref(c))
    # CMolDynConfig::resetToDefaults (MolDynConfig.cpp:307-319)
    assert (c.xyz_file, c.dcd_file, c.xtc_file) == (b"traj.xyz", b"traj.dcd", b"traj.xtc")
    assert (c.pdistr_file, c.vdistr_file) == (b"distr", b"distr")
    assert (c.include_xyz, c.include_dcd, c.include_xtc, c.include_pdistr, c.include_vdistr) == (0, 1, 0, 0, 0)
    assert c.xtc_precision == 1000.0 and c.max_pdistr == 0.3 and c.max_vdistr == 100.0e3
    m = config.MolDynConfig()
    err = C.create_string_buffer(256)
    for key, value, field, expect in [("xyzfile", "a.xyz", "xyz_file", b"a.xyz"),
                                      ("dcdfile", "b.dcd", "dcd_file", b"b.dcd"),
                                      ("pdistrfile", "pd", "pdistr_file", b"pd"),
                                      ("vdistrfile", "vd", "vdistr_file", b"vd"),
                                      ("includexyzfile", "yes", "include_xyz", 1),
                                      ("includedcdfile", "no", "include_dcd", 0),
                                      ("includepdistrfile", "yes", "include_pdistr", 1),
                                      ("includevdistrfile", "yes", "include_vdistr", 1),
                                      ("maxpdistr", "0.5", "max_pdistr", 0.5),
                                      ("maxvdistr", "2e4", "max_vdistr", 2e4),
                                      ("xtcprecision", "100", "xtc_precision", 100.0)]:
        assert lib.mtmd_config_set(C.byref(c), key.encode(), value.encode(), err, 256) == 0, key
        assert getattr(c, field) == expect, key
        assert m.set(key, value) == ""
        assert getattr(m.cfg, field) == expect, key
    for key in ("includexyzfile", "includextcfile", "includepdistrfile", "includevdistrfile", "verboseoutput"):
        msg = f"Error: '{key}' parameter should be either 'yes' or 'no'!"
        assert lib.mtmd_config_set(C.byref(c), key.encode(), b"y", err, 256) == abi.MTMD_EINVAL
        assert err.value.decode() == msg
        assert m.set(key, "y") == msg
