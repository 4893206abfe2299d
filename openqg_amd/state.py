"""QG restart files: `save_qg` and `restart_qg` (src/state.F90:29-44,82-112).

A file holds what the run loop needs to resume a basin: the two pressure time
levels.  `save_qg` writes, as src/state.F90:92-110 does through
src/util/ncutils.F90:

- a netCDF classic file (nf90_create with NF90_CLOBBER, ncutils.F90:110-121),
  no global attributes;
- dimensions ``nxp``, ``nyp``, ``nl``;
- ``x(nxp)`` and ``y(nyp)``: the p-grid coordinates in km
  (``1.0d-3 * xp``, src/util/units.F90:17-21; xp and yp from
  src/glam/box.F90:143,154), each with ``units = "km"``;
- ``p`` and ``pm``: Fortran ``(nxp, nyp, nl)``, which is ``(nl, nyp, nxp)`` in
  the file's C order and in this package's arrays, with ``units`` set to
  ``"Current pressure"`` and ``"Previous pressure"`` (nc_def_double passes its
  fourth argument to the ``units`` attribute, ncutils.F90:264-266).

The run loop names them ``<basin>_qg_restart.nc`` (every ``noutre`` steps),
``<basin>_qg_lastday.nc`` (end of run) and ``<basin>_qg_invalid.nc`` (when the
state stops being finite), src/openqg.F90:499-512,529-532.  A restart reads
``p`` and ``pm`` back and then derives q, qm and the constraint scalars from
them (init_pv, init_constraint, init_mass_constr; src/openqg.F90:237-243),
which is ``QG.init_state``.

The file format is written and read with scipy's netCDF classic module, an
implementation of the format independent of this package.
"""
from __future__ import annotations

import os

import numpy as np
from scipy.io import netcdf_file

_UNITS = {"p": "Current pressure", "pm": "Previous pressure"}


def write_qg_file(path: str, box, p, pm) -> None:
    """Write p and pm of a basin on `box` (a `Box`) to the restart file `path`."""
    shape = (box.nl, box.nyp, box.nxp)
    fields = {}
    for name, a in (("p", p), ("pm", pm)):
        a = np.asarray(a, dtype=np.float64)
        if a.shape != shape:
            raise ValueError(f"{name}: shape {a.shape} != {shape}")
        fields[name] = a
    with netcdf_file(path, "w", version=1) as f:
        f.createDimension("nxp", box.nxp)
        f.createDimension("nyp", box.nyp)
        f.createDimension("nl", box.nl)
        for name, dim, vals in (("x", "nxp", box.xp), ("y", "nyp", box.yp)):
            v = f.createVariable(name, "d", (dim,))
            v.units = "km"
            v[:] = 1.0e-3 * vals
        for name in ("p", "pm"):
            v = f.createVariable(name, "d", ("nl", "nyp", "nxp"))
            v.units = _UNITS[name]
            v[:] = fields[name]


def read_qg_file(path: str, box):
    """(p, pm) of the restart file `path`, checked against the grid of `box`.

    Raises ValueError when a variable is missing or its dimensions differ from
    (nl, nyp, nxp) of `box` (nc_get_double stops the reference in both cases,
    src/util/ncutils.F90:194-208)."""
    shape = (box.nl, box.nyp, box.nxp)
    out = []
    with netcdf_file(path, "r", mmap=False) as f:
        for name in ("p", "pm"):
            if name not in f.variables:
                raise ValueError(f"{path}: no variable '{name}'")
            v = f.variables[name]
            if tuple(v.shape) != shape:
                raise ValueError(f"{path}: '{name}' has shape {tuple(v.shape)}, the basin needs {shape}")
            out.append(np.array(v[:], dtype=np.float64))
    return out[0], out[1]


def save_qg(outdir: str, filename: str, qg) -> str:
    """save_qg (src/state.F90:82-112): write p and pm of `qg` (a `QG`) to
    outdir/filename; returns the path."""
    _, _, p, pm = qg.get_state()
    path = os.path.join(outdir, filename)
    write_qg_file(path, qg.b, p, pm)
    return path


def restart_qg(indir: str, filename: str, qg, init: bool = True) -> None:
    """restart_qg (src/state.F90:29-44): load p and pm of `qg` from
    indir/filename.  With `init` (the default) q, qm and the constraint
    scalars are then derived from them as src/openqg.F90:237-243 does."""
    p, pm = read_qg_file(os.path.join(indir, filename), qg.b)
    qg.set_state(p=p, pm=pm)
    if init:
        qg.init_state()


# ---------------------------------------------------------------- mixed layers
# save_oml / save_aml (src/state.F90:114-176) and restart_oml / restart_aml
# (:46-80): dimensions nxt, nyt, x(nxt) and y(nyt) at T points in km, and the
# fields (nyt, nxt) in the file's C order.
_ML_UNITS = {"sst": "Current mixed layer temp", "sstm": "Previous mixed layer temp",
             "ast": "Current mixed layer temp", "astm": "Previous mixed layer temp",
             "hmixa": "Current mixed layer height", "hmixam": "Previous mixed layer height"}


def write_ml_file(path: str, box, fields: dict) -> None:
    shape = (box.nyt, box.nxt)
    with netcdf_file(path, "w", version=1) as f:
        f.createDimension("nxt", box.nxt)
        f.createDimension("nyt", box.nyt)
        xt = box.xp[:-1] + 0.5 * box.dx
        yt = box.yp[:-1] + 0.5 * box.dy
        for name, dim, vals in (("x", "nxt", xt), ("y", "nyt", yt)):
            v = f.createVariable(name, "d", (dim,))
            v.units = "km"
            v[:] = 1.0e-3 * vals
        for name, a in fields.items():
            a = np.asarray(a, dtype=np.float64)
            if a.shape != shape:
                raise ValueError(f"{name}: shape {a.shape} != {shape}")
            v = f.createVariable(name, "d", ("nyt", "nxt"))
            v.units = _ML_UNITS[name]
            v[:] = a


def read_ml_file(path: str, box, names):
    shape = (box.nyt, box.nxt)
    out = []
    with netcdf_file(path, "r", mmap=False) as f:
        for name in names:
            if name not in f.variables:
                raise ValueError(f"{path}: no variable '{name}'")
            v = f.variables[name]
            if tuple(v.shape) != shape:
                raise ValueError(f"{path}: '{name}' has shape {tuple(v.shape)}, the mixed layer needs {shape}")
            out.append(np.array(v[:], dtype=np.float64))
    return out


def save_oml(outdir: str, filename: str, oml) -> str:
    """save_oml (src/state.F90:114-143) of an OceanMixedLayer."""
    sst, sstm = oml.get_state()
    path = os.path.join(outdir, filename)
    write_ml_file(path, oml.qg.b, {"sst": sst, "sstm": sstm})
    return path


def restart_oml(indir: str, filename: str, oml) -> None:
    """restart_oml (src/state.F90:46-61)."""
    sst, sstm = read_ml_file(os.path.join(indir, filename), oml.qg.b, ("sst", "sstm"))
    oml.set_state(sst, sstm)


def save_aml(outdir: str, filename: str, cpl) -> str:
    """save_aml (src/state.F90:145-176) of the atmospheric mixed layer a Coupler holds."""
    ast, astm, hmixa, hmixam = cpl.get_aml_state()
    path = os.path.join(outdir, filename)
    write_ml_file(path, cpl.atm.b, {"ast": ast, "astm": astm, "hmixa": hmixa, "hmixam": hmixam})
    return path


def restart_aml(indir: str, filename: str, cpl) -> None:
    """restart_aml (src/state.F90:63-80)."""
    f = read_ml_file(os.path.join(indir, filename), cpl.atm.b, ("ast", "astm", "hmixa", "hmixam"))
    cpl.set_aml_state(*f)
