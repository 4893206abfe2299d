"""Host mirror of OpenQG's ocean mixed layer over the HIP library.

`Temp` carries temp_type (src/mixed/mixed.F90:28-43) and `zonal_temp`
(mixed.F90:145-152); `OceanMixedLayer` is ocean_mixed_type
(src/mixed/omlsubs.F90:13-20) bound to one ocean basin (`QG`): `init_ocean_ml`
mirrors omlsubs.F90:27-67, `step` is step_oml (omlsubs.F90:69-119) and `run`
the ocean half of dynamic_step (driver.F90:73-81: step_ocn, then
suppress_ocean).  The mixed-layer temperature lives on the GPU next to the
basin's state: a step reads layer 1 of the basin's current pressure and its
resident wind stress (set by `QG.set_wind`), and writes the layer 1/2
entrainment into the basin's resident forcing, which the basin's next step
reads.  T fields are numpy arrays of shape (nyt, nxt) (the C view of the
Fortran (nxt, nyt) arrays).
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

import numpy as np

from ._lib import OQG_HOST, OqgOmlDesc
from .qg import QG, Box, _as_f64_contig


@dataclass
class Temp:
    """temp_type: the vertical and zonal temperature profile of a basin."""

    T1_rel: float   # T_bar - T(1)
    dT_12: float    # T(1) - T(2)
    dT_NS: float    # (T(N) - T(S)) / 2
    y0: float
    ymax: float
    rho_cp: float


def init_temp(b: Box, T1_rel: float, T2_rel: float, dT_NS: float, rho: float, cp: float) -> Temp:
    """init_temp (mixed.F90:125-143)."""
    return Temp(T1_rel=T1_rel, dT_12=T1_rel - T2_rel, dT_NS=dT_NS, y0=float(b.yp[0]) + 0.5 * b.yl,
                ymax=float(b.yp[0]) + b.yl, rho_cp=rho * cp)


def init_temp_from_rad(b: Box, ga: Box, T1: float, T2: float, Tbar: float, rbtm: float, fspco: float, rho: float,
                       cp: float) -> Temp:
    """init_temp_from_rad (mixed.F90:102-123): the zonal profile spans the atmosphere's box."""
    return Temp(T1_rel=T1 - Tbar, dT_12=T1 - T2, dT_NS=abs(rbtm * fspco), y0=float(ga.yp[0]) + 0.5 * ga.yl,
                ymax=float(ga.yp[0]) + ga.yl, rho_cp=rho * cp)


def sin_lat(b: Box, y: float, y0: float, ymax: float, bar: float) -> float:
    """sin_lat (glam/numerics.F90:516-526)."""
    return -0.5 * math.copysign(abs(bar), b.fnot) * math.sin((math.pi / 2.0) * (y - y0) / (ymax - y0))


def zonal_temp(b: Box, temp: Temp, j: int) -> float:
    """zonal_temp (mixed.F90:145-152) at T row j (1-based); yt(j) = yp(j) + dy/2 (box.F90:146-150)."""
    return sin_lat(b, float(b.yp[j - 1]) + 0.5 * b.dy, temp.y0, temp.ymax, temp.dT_NS)


class OceanMixedLayer:
    """ocean_mixed_type on the device, bound to the ocean basin `qg`."""

    def __init__(self, qg: QG, temp: Temp, fluxn: bool, fluxs: bool, st2d: float, st4d: float, ycexp: float,
                 hm: float | None = None):
        """init_ocean_ml (omlsubs.F90:27-67): the flux-wall values are zonal_temp at the
        outer T rows; a heat-flux wall on the pole-ward side of the hemisphere is rejected."""
        b = qg.b
        hm = float(b.hm if hm is None else hm)
        if fluxs and b.fnot < 0.0:
            raise ValueError("southern boundary heat flux in a southern-hemisphere basin (omlsubs.F90:43-53)")
        if fluxn and b.fnot > 0.0:
            raise ValueError("northern boundary heat flux in a northern-hemisphere basin (omlsubs.F90:54-64)")
        self.qg, self.temp = qg, temp
        self.fluxn, self.fluxs = bool(fluxn), bool(fluxs)
        self.st2d, self.st4d, self.ycexp, self.hm = float(st2d), float(st4d), float(ycexp), hm
        self.nval, self.sval = zonal_temp(b, temp, b.nyt), zonal_temp(b, temp, 1)
        d = OqgOmlDesc()
        d.hm, d.T1_rel, d.dT_12, d.rho_cp = hm, temp.T1_rel, temp.dT_12, temp.rho_cp
        d.st2d, d.st4d, d.ycexp, d.nval, d.sval = self.st2d, self.st4d, self.ycexp, self.nval, self.sval
        d.fluxn, d.fluxs = int(self.fluxn), int(self.fluxs)
        self._lib = qg._lib
        h = ctypes.c_void_p()
        rc = self._lib.oqg_oml_create(qg._h, ctypes.byref(d), ctypes.byref(h))
        if rc != 0:
            msg = self._lib.oqg_oml_last_error(None)
            raise RuntimeError(f"oqg_oml_create failed (code {rc}): {msg.decode() if msg else ''}")
        self._h = h
        qg._attach(self)   # closed before the basin it reads and writes

    @property
    def shape(self):
        return (self.qg.b.nyt, self.qg.b.nxt)

    def params(self) -> dict:
        """The parameters in the oracle's form (hm, T1_rel, dT_12, rho_cp, d2, d4, ycexp, nval, sval, fluxn, fluxs)."""
        return dict(hm=self.hm, T1_rel=self.temp.T1_rel, dT_12=self.temp.dT_12, rho_cp=self.temp.rho_cp, d2=self.st2d,
                    d4=self.st4d, ycexp=self.ycexp, nval=self.nval, sval=self.sval, fluxn=self.fluxn,
                    fluxs=self.fluxs)

    # ------------------------------------------------------------ lifecycle
    def close(self):
        if getattr(self, "_h", None) is not None:
            self._lib.oqg_oml_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != 0:
            msg = self._lib.oqg_oml_last_error(self._h)
            raise RuntimeError(f"{what} failed (code {rc}): {msg.decode() if msg else ''}")

    # ------------------------------------------------------------ state
    def set_state(self, sst=None, sstm=None):
        objs = [_as_f64_contig(a) for a in (sst, sstm)]
        kinds = {o[2] for o in objs if o[0] is not None}
        if len(kinds) > 1:
            raise ValueError("pass both fields on the host or both on the device")
        for o in objs:
            if o[0] is not None and tuple(o[0].shape) != self.shape:
                raise ValueError(f"T field shape {tuple(o[0].shape)} != {self.shape}")
        kind = kinds.pop() if kinds else OQG_HOST
        self._check(self._lib.oqg_oml_set_state(self._h, objs[0][1], objs[1][1], kind), "oqg_oml_set_state")

    def get_state(self):
        """(sst, sstm) as numpy arrays (nyt, nxt)."""
        sst, sstm = np.empty(self.shape), np.empty(self.shape)
        self._check(self._lib.oqg_oml_get_state(self._h, sst.ctypes.data_as(ctypes.c_void_p),
                                                sstm.ctypes.data_as(ctypes.c_void_p), OQG_HOST), "oqg_oml_get_state")
        return sst, sstm

    def set_fnet(self, fnet):
        """The net surface heat flux into the ocean at T points (ocn_cpl%fnet)."""
        o = _as_f64_contig(fnet)
        if tuple(o[0].shape) != self.shape:
            raise ValueError(f"fnet shape {tuple(o[0].shape)} != {self.shape}")
        self._check(self._lib.oqg_oml_set_fnet(self._h, o[1], o[2]), "oqg_oml_set_fnet")

    # ------------------------------------------------------------ time stepping
    def step(self, tdto: float, compute_ent: bool = True):
        """step_oml (omlsubs.F90:69-119) on the basin's current layer-1 pressure."""
        self._check(self._lib.oqg_oml_step(self._h, float(tdto), int(bool(compute_ent))), "oqg_oml_step")

    def suppress(self):
        """The mixed-layer half of suppress_ocean (driver.F90:165-167)."""
        self._check(self._lib.oqg_oml_suppress(self._h), "oqg_oml_suppress")

    def run(self, tdto: float, nsteps: int, step0: int = 0, supp_every: int = 25, compute_ent: bool = True):
        """nsteps of step_ocn (driver.F90:89-110) with suppress_ocean every supp_every
        steps (driver.F90:78-81), replayed from captured graphs."""
        self._check(self._lib.oqg_run_ocean(self._h, float(tdto), int(nsteps), int(step0), int(supp_every),
                                            int(bool(compute_ent))), "oqg_run_ocean")
