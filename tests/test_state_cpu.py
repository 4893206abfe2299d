"""Restart files (openqg_amd/state.py) against src/state.F90:82-112.

The layout is checked by a parser of the netCDF classic format written here
from the published format specification (magic "CDF\\x01", dimension,
attribute and variable lists, big-endian values at each variable's `begin`
offset), so a writer that transposed the fields, wrote them little-endian or
named a dimension differently fails independently of the library that wrote
the file.
"""
import os
import struct

import numpy as np
import pytest

from openqg_amd.qg import Box
from openqg_amd.state import read_qg_file, write_qg_file

NC_DIMENSION, NC_VARIABLE, NC_ATTRIBUTE = 0x0A, 0x0B, 0x0C
NC_CHAR, NC_DOUBLE = 2, 6
_SIZE = {1: 1, 2: 1, 3: 2, 4: 4, 5: 4, 6: 8}


def _parse_classic(buf: bytes):
    """Header of a netCDF classic (version 1) file: dims, variables with
    their dimension names, type, attributes and data offset."""
    assert buf[:4] == b"CDF\x01"
    pos = 8   # magic, numrecs

    def i32():
        nonlocal pos
        (v,) = struct.unpack(">i", buf[pos:pos + 4])
        pos += 4
        return v

    def name():
        nonlocal pos
        n = i32()
        s = buf[pos:pos + n].decode()
        pos += (n + 3) // 4 * 4
        return s

    def atts():
        nonlocal pos
        tag, n = i32(), i32()
        assert tag in (0, NC_ATTRIBUTE)
        out = {}
        for _ in range(n):
            k, t, m = name(), i32(), i32()
            raw = buf[pos:pos + m * _SIZE[t]]
            pos += (m * _SIZE[t] + 3) // 4 * 4
            out[k] = (t, raw.decode() if t == NC_CHAR else raw)
        return out

    tag, n = i32(), i32()
    assert tag in (0, NC_DIMENSION)
    dims = [(name(), i32()) for _ in range(n)]
    gatts = atts()
    tag, n = i32(), i32()
    assert tag == NC_VARIABLE
    vars_ = {}
    for _ in range(n):
        vn = name()
        ids = [i32() for _ in range(i32())]
        va = atts()
        t, vsize, begin = i32(), i32(), i32()
        vars_[vn] = dict(dims=[dims[i][0] for i in ids], shape=[dims[i][1] for i in ids], type=t,
                         atts=va, begin=begin, vsize=vsize)
    return dims, gatts, vars_


def _box():
    return Box(11, 7, 3, 30.0e3, 25.0e3, 1.0e-4, 2.0e-11, [300.0, 1100.0, 2600.0], x0=60.0e3, y0=30.0e3)


def _fields(b, seed=0):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal((b.nl, b.nyp, b.nxp)) * 1e3, rng.standard_normal((b.nl, b.nyp, b.nxp)) * 1e3)


def test_restart_file_layout_matches_save_qg(tmp_path):
    b = _box()
    p, pm = _fields(b)
    path = os.path.join(tmp_path, "ocean_qg_lastday.nc")
    write_qg_file(path, b, p, pm)
    buf = open(path, "rb").read()
    dims, gatts, v = _parse_classic(buf)
    assert dims == [("nxp", b.nxp), ("nyp", b.nyp), ("nl", b.nl)]   # state.F90:94-96
    assert gatts == {}
    assert list(v) == ["x", "y", "p", "pm"]                         # state.F90:98-101
    assert v["x"]["dims"] == ["nxp"] and v["y"]["dims"] == ["nyp"]
    # Fortran (xdim, ydim, zdim) is (nl, nyp, nxp) in the file's C order
    assert v["p"]["dims"] == ["nl", "nyp", "nxp"] and v["pm"]["dims"] == ["nl", "nyp", "nxp"]
    for k, units in (("x", "km"), ("y", "km"), ("p", "Current pressure"), ("pm", "Previous pressure")):
        assert v[k]["type"] == NC_DOUBLE
        assert v[k]["atts"] == {"units": (NC_CHAR, units)}

    def data(k):
        n = int(np.prod(v[k]["shape"]))
        return np.frombuffer(buf, dtype=">f8", count=n, offset=v[k]["begin"]).reshape(v[k]["shape"])

    # x(i) = 1e-3 * (x0 + (i-1) dx) km (box.F90:154, units.F90:17-21)
    np.testing.assert_array_equal(data("x"), 1.0e-3 * (60.0e3 + np.arange(11) * 30.0e3))
    np.testing.assert_array_equal(data("y"), 1.0e-3 * (30.0e3 + np.arange(7) * 25.0e3))
    np.testing.assert_array_equal(data("p"), p)
    np.testing.assert_array_equal(data("pm"), pm)
    # the Fortran view: p(i, j, k) at C index [k-1, j-1, i-1]
    assert data("p")[2, 1, 4] == p[2, 1, 4]


def test_restart_round_trip_is_bitwise(tmp_path):
    b = _box()
    p, pm = _fields(b, 1)
    p[0, 0, 0], pm[1, 2, 3] = np.finfo(np.float64).tiny, -0.0
    path = os.path.join(tmp_path, "atmos_qg_restart.nc")
    write_qg_file(path, b, p, pm)
    rp, rpm = read_qg_file(path, b)
    assert rp.dtype == np.float64 and rp.flags.c_contiguous
    assert rp.tobytes() == p.tobytes() and rpm.tobytes() == pm.tobytes()


def test_restart_rejects_other_grids_and_missing_fields(tmp_path):
    b = _box()
    p, pm = _fields(b)
    path = os.path.join(tmp_path, "x_qg_lastday.nc")
    write_qg_file(path, b, p, pm)
    other = Box(11, 7, 2, 30.0e3, 25.0e3, 1.0e-4, 2.0e-11, [1000.0, 3000.0])
    with pytest.raises(ValueError, match="shape"):
        read_qg_file(path, other)
    with pytest.raises(ValueError, match="shape"):
        write_qg_file(path, b, p[:2], pm)
    from scipy.io import netcdf_file

    bad = os.path.join(tmp_path, "nopm.nc")
    with netcdf_file(bad, "w") as f:
        f.createDimension("nxp", 11)
        f.createDimension("nyp", 7)
        f.createDimension("nl", 3)
        f.createVariable("p", "d", ("nl", "nyp", "nxp"))[:] = p
    with pytest.raises(ValueError, match="no variable 'pm'"):
        read_qg_file(bad, b)
