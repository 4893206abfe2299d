"""Host mirror of OpenQG's ocean-atmosphere coupling over the HIP library.

`Coupler` binds an ocean basin with its mixed layer (`OceanMixedLayer`) to an
atmosphere channel (`QG`) and holds the atmospheric mixed layer
(atmos_mixed_type, src/mixed/amlsubs.F90:30-53).  On the device it computes the
coupled wind stress (xforc, src/windstress/windstress.F90:102-316, with the
bicubic interpolation of src/glam/bicubic.F90), the heat-flux forcing
(compute_forcing, src/mixed/forcing.F90:18-99) and step_aml
(amlsubs.F90:184-326), and `run` is the time loop of dynamic_step
(src/driver.F90:36-87, src/openqg.F90:483-488) replayed from HIP graphs.

`load_grid` mirrors src/glam/grid.F90:37-58.  The radiation coefficients
(radiate_type, src/mixed/radsubs.F90:14-53) and the entrainment coefficients
(init_entrain_coeff, amlsubs.F90:93-129) are inputs, in `params`.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from ._lib import OQG_CPL_ATMOS_ENT, OQG_CPL_OCEAN_ENT, OQG_HOST, OqgCplDesc, _CPL_D
from .mixed import OceanMixedLayer
from .qg import QG, _as_f64_contig


@dataclass
class Grid:
    """grid_type (grid.F90:11-27)."""

    ndxr: int
    nxtoar: int
    nytoar: int
    nxpoar: int
    nypoar: int
    nxaooc: int
    nyaooc: int
    nx1: int
    ny1: int


def load_grid(go, ga) -> Grid:
    """load_grid (grid.F90:37-58) for boxes with nxp, nyp, dx."""
    ndxr = int(np.floor(ga.dx / go.dx + 0.5))   # nint of a positive ratio
    nxto, nyto, nyta = go.nxp - 1, go.nyp - 1, ga.nyp - 1
    nxtoar, nytoar = nxto // ndxr, nyto // ndxr
    return Grid(ndxr, nxtoar, nytoar, nxtoar + 1, nytoar + 1, nxto // ndxr, nyto // ndxr, 1,
                1 + (nyta - nyto // ndxr) // 2)


class Coupler:
    """The coupled ocean (`ocean_ml` and its basin) and atmosphere (`atmos`).

    `params` holds: xo0, yo0, xa0, ya0 (box origins), hmo, hma, rho_o, rho_a, cdat,
    tau_udiff, the radiation coefficients D0up, Dmdown, Dup0, xlamda, fspco, Aup00,
    Adown10, Adown11, the atmospheric mixed layer's hmamin, hmadmp, xcexp, T1_rel,
    dT_12, rho_cp, at2d, at4d, ahmd, and the entrainment coefficients aface (nl-1),
    bface, dface.  `xc1ast` (nyta, nxta) and `dtop` (nypa, nxpa) default to zero."""

    def __init__(self, ocean_ml: OceanMixedLayer, atmos: QG, params: dict, xc1ast=None, dtop=None):
        self.oml, self.atm, self.ocn = ocean_ml, atmos, ocean_ml.qg
        self.params = dict(params)
        self.grid = load_grid(self.ocn.b, atmos.b)
        d = OqgCplDesc()
        for n in ("xo0", "yo0", "xa0", "ya0", "hmo", "hma", "rho_o", "rho_a", "cdat") + _CPL_D:
            setattr(d, n, float(params[n]))
        d.tau_udiff = int(bool(params.get("tau_udiff", False)))
        for k, v in enumerate(params["aface"]):
            d.aface[k] = float(v)
        d.bface, d.dface = float(params["bface"]), float(params["dface"])
        keep = []
        nyta, nxta = atmos.b.nyp - 1, atmos.b.nxp - 1
        for name, arr, shape in (("xc1ast", xc1ast, (nyta, nxta)), ("dtop", dtop, (atmos.b.nyp, atmos.b.nxp))):
            if arr is not None:
                a = np.ascontiguousarray(arr, dtype=np.float64)
                if a.shape != shape:
                    raise ValueError(f"{name} shape {a.shape} != {shape}")
                keep.append(a)
                setattr(d, name, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        self._lib = atmos._lib
        h = ctypes.c_void_p()
        rc = self._lib.oqg_cpl_create(ocean_ml._h, atmos._h, ctypes.byref(d), ctypes.byref(h))
        if rc != 0:
            msg = self._lib.oqg_cpl_last_error(None)
            raise RuntimeError(f"oqg_cpl_create failed (code {rc}): {msg.decode() if msg else ''}")
        self._h = h
        # closed before either basin or the ocean mixed layer it reads and writes
        self.ocn._attach(self)
        atmos._attach(self)

    # ------------------------------------------------------------ lifecycle
    def close(self):
        if getattr(self, "_h", None) is not None:
            self._lib.oqg_cpl_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != 0:
            msg = self._lib.oqg_cpl_last_error(self._h)
            raise RuntimeError(f"{what} failed (code {rc}): {msg.decode() if msg else ''}")

    @property
    def tshape_a(self):
        return (self.atm.b.nyp - 1, self.atm.b.nxp - 1)

    # ------------------------------------------------------------ state
    def set_aml_state(self, ast=None, astm=None, hmixa=None, hmixam=None):
        objs = [_as_f64_contig(a) for a in (ast, astm, hmixa, hmixam)]
        kinds = {o[2] for o in objs if o[0] is not None}
        if len(kinds) > 1:
            raise ValueError("pass all fields on the host or all on the device")
        for o in objs:
            if o[0] is not None and tuple(o[0].shape) != self.tshape_a:
                raise ValueError(f"T field shape {tuple(o[0].shape)} != {self.tshape_a}")
        kind = kinds.pop() if kinds else OQG_HOST
        self._check(self._lib.oqg_cpl_set_aml_state(self._h, *(o[1] for o in objs), kind), "oqg_cpl_set_aml_state")

    def get_aml_state(self):
        """(ast, astm, hmixa, hmixam), each (nyta, nxta)."""
        outs = [np.empty(self.tshape_a) for _ in range(4)]
        self._check(self._lib.oqg_cpl_get_aml_state(self._h, *(o.ctypes.data_as(ctypes.c_void_p) for o in outs),
                                                    OQG_HOST), "oqg_cpl_get_aml_state")
        return tuple(outs)

    def get_fnet(self):
        """(fnetoc (nyto, nxto), fnetat (nyta, nxta)) of the last compute_forcing."""
        fo = np.empty((self.ocn.b.nyp - 1, self.ocn.b.nxp - 1))
        fa = np.empty(self.tshape_a)
        self._check(self._lib.oqg_cpl_get_fnet(self._h, fo.ctypes.data_as(ctypes.c_void_p),
                                               fa.ctypes.data_as(ctypes.c_void_p), OQG_HOST), "oqg_cpl_get_fnet")
        return fo, fa

    # ------------------------------------------------------------ the exchange
    def xforc(self):
        """xforc and ekman_from_tau for both basins (driver.F90:47-53)."""
        self._check(self._lib.oqg_cpl_xforc(self._h), "oqg_cpl_xforc")

    def forcing(self):
        """compute_forcing (forcing.F90:18-99)."""
        self._check(self._lib.oqg_cpl_forcing(self._h), "oqg_cpl_forcing")

    def aml_step(self, tdta: float, compute_ent: bool = True):
        """step_aml (amlsubs.F90:184-229)."""
        self._check(self._lib.oqg_cpl_aml_step(self._h, float(tdta), int(bool(compute_ent))), "oqg_cpl_aml_step")

    def run(self, nsteps: int, step0: int, tdto: float, tdta: float, nstr: int, osupp: int | None = None,
            asupp: int = 100, ocean_ent: bool = True, atmos_ent: bool = True):
        """nsteps atmospheric steps of dynamic_step from step0 (openqg.F90:483-488); the
        ocean steps every nstr, is suppressed every osupp (default 25 nstr) and the
        atmosphere every asupp steps."""
        osupp = 25 * nstr if osupp is None else osupp
        flags = (OQG_CPL_OCEAN_ENT if ocean_ent else 0) | (OQG_CPL_ATMOS_ENT if atmos_ent else 0)
        self._check(self._lib.oqg_cpl_run(self._h, int(nsteps), int(step0), float(tdto), float(tdta), int(nstr),
                                          int(osupp), int(asupp), flags), "oqg_cpl_run")
